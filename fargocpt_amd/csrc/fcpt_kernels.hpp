// Launchers of the gfx950 hydro-step kernels (fcpt_kernels.hip).  Every
// launcher enqueues on the given stream and reads step-varying scalars (dt,
// time, planets) from device memory, so a sequence of launchers can be
// captured into a HIP graph and replayed.
#pragma once

#include <vector>

#include "fcpt_device.hpp"

namespace fcpt {

void launch_ring_mean(const KParams &P, hipStream_t st, int mode);
void launch_cfl(const KParams &P, hipStream_t st, int mode);
void launch_set_dt(const KParams &P, hipStream_t st, double dt);
void launch_set_clock(const KParams &P, hipStream_t st, double t);
void launch_set_time(const KParams &P, hipStream_t st, double t, double last_dt, int n_monitor);
void launch_set_params(const KParams &P, hipStream_t st, const fcpt_step_params &sp,
                       const BodySmoothing &sm);
void launch_potential(const KParams &P, hipStream_t st);
void launch_sources(const KParams &P, hipStream_t st);
bool launch_artificial_viscosity(const KParams &P, hipStream_t st, int with_viscosity);
void launch_recalc_viscosity(const KParams &P, hipStream_t st);
void launch_stress(const KParams &P, hipStream_t st);
void launch_visc_update(const KParams &P, hipStream_t st);
void launch_substep3(const KParams &P, hipStream_t st);
void launch_heatcool_cfl(const KParams &P, hipStream_t st);
void launch_boundaries(const KParams &P, hipStream_t st, int final_call, int advance_time);
void launch_transport(const KParams &P, hipStream_t st);
void launch_set_mid_params(const KParams &P, hipStream_t st, const fcpt_step_params &sp,
                           const BodySmoothing &sm);
void launch_leapfrog_phase(const KParams &P, hipStream_t st, int phase);
void launch_pressure(const KParams &P, hipStream_t st);
void launch_final_boundaries_derived(const KParams &P, hipStream_t st, int advance_time);
void launch_copy_velocities(const KParams &P, hipStream_t st);
void launch_fill_rows(const KParams &P, hipStream_t st, void *dst, const double *ring,
                      const double *bg, int rows);
void launch_transport_to(const KParams &P, const KParams &out, hipStream_t st);
// the boundary call before the transport (ghost rings) and the transport, the
// ghost rings and the FARGO ring mean in one launch; final_advance > 0: the
// velocity kernel also runs the final boundary call (damping, ghost rings,
// the clock with that advance_time) and the derived fields (Euler)
void launch_transport_ghost(const KParams &P, const KParams &out, hipStream_t st, int final_advance);
// the fused step kernels write the velocities of view P to P.valt; the
// swapped view reads those buffers and writes the fields
KParams swapped_view(const KParams &P);
void launch_src_fused(const KParams &P, hipStream_t st);
void launch_src_energy(const KParams &P, hipStream_t st);
void launch_av_fused(const KParams &P, hipStream_t st, int with_viscosity);
void launch_visc_fused(const KParams &P, hipStream_t st);
// the fused viscous pass followed by SubStep3 (adiabatic disks); one walk
// that does not store the stress tensor where the grid is large enough
void launch_visc_energy(const KParams &P, hipStream_t st);
// the kick's potential pass, then its source terms and artificial viscosity as
// one walk (TW or SN artificial viscosity on a grid large enough for it),
// velocities to the view's partner buffers; then launch_visc_energy
// (adiabatic) or launch_visc_fused on the swapped view
bool kick_walk_applies(const KParams &P);
// the rings per thread the walks use for this disk (1: the cell-parallel form)
int resolve_kick_tile(const KParams &P);
int resolve_visc_tile(const KParams &P);
// wavefronts of one kick-walk or viscous-walk launch of a disk at `tile`
// rings per thread (the mean of the two walks' segment counts)
long walk_waves(const KParams &P, int tile);
void launch_kick(const KParams &P, hipStream_t st, int fold_mode);
// the potential; fold_mode >= 0: in the same launch, the CFL fold of
// k_cfl_rings' slots (launch_cfl_cells) with that mode
void launch_potential_fold(const KParams &P, hipStream_t st, int fold_mode);
void launch_radial_sweep(const KParams &P, hipStream_t st);
void launch_azimuthal_sweep(const KParams &P, hipStream_t st);
void launch_cfl_cells(const KParams &P, hipStream_t st);
void launch_copy4(const KParams &P, hipStream_t st);
void launch_derived(const KParams &P, hipStream_t st);
void launch_iso_profile(const KParams &P, hipStream_t st);
// the locally isothermal ring profiles of c_s, H and nu (geometry and parameters only)
void launch_iso_rings(const KParams &P, hipStream_t st);
// KParams::ring_tab from the metric arrays (once, at context creation)
void launch_ring_tables(const KParams &P, hipStream_t st);
void launch_totals(const KParams &P, hipStream_t st, double radius, double *ring_scratch,
                   double *out2);
// split storage: deviations of `field` (rows x nphi floats) re-expressed from
// background oldbg to newbg (both device arrays)
void launch_rebase(const KParams &P, hipStream_t st, void *field, const double *oldbg,
                   const double *newbg, int rows);
// split storage: out[i] = ring mean of the field's row i (rows rows)
void launch_split_ring_mean(const KParams &P, hipStream_t st, const void *field, const double *bg,
                            int rows, double *out);
// bytes of one stored value of a 2-D field (8 in FP64 storage, 4 in split storage)
size_t element_size(const KParams &P);
// bytes per value of one field: the potential is FP64 in both storage modes
size_t field_element_size(const KParams &P, int field);
// rings per thread of the radial sweep (requested > 0 overrides the size-based choice)
int radial_tile_for(int nr, int ns, int requested);
// fills the launch-shape fields of a disk's parameters (the CFL grid)
void init_launch_fields(KParams &P);

// The disk's gravity on up to kForceMaxBodies bodies (ComputeDiskOnPlanetAccel)
constexpr int kForceMaxBodies = FCPT_MAX_PLANETS;
struct ForceBodies {
    int n, flags;
    double x[kForceMaxBodies], y[kForceMaxBodies], r[kForceMaxBodies], r_cubic[kForceMaxBodies];
    double smooth_fixed[kForceMaxBodies];  // >= 0: that smoothing length, < 0: thickness x H(cell)
};
// sig1d: nr doubles (ring means of Sigma), part: kCflMaxBlocks x 4 kForceMaxBodies
// doubles, out4: 4 kForceMaxBodies doubles (inner x, inner y, outer x, outer y per body)
void launch_disk_force(const KParams &P, hipStream_t st, const ForceBodies &B, double *sig1d,
                       double *part, double *out4);

// Batched step: one work-table entry per block of one disk's grid
struct BatchEntry {
    unsigned disk, yx;  // yx = by << 16 | bx
};
enum BatchKind {
    B_RING_MEAN, B_CFL_CELLS, B_CFL_FINAL, B_POTENTIAL, B_SRC_VR, B_SRC_VPHI,
    B_SRC_ENERGY, B_TW_Q, B_TW_VEL, B_SN_Q, B_SN_VEL, B_TEMP_FLOOR, B_RECALC_VISC, B_STRESS,
    B_VISC_UPDATE, B_SUBSTEP3, B_GHOST, B_RADIAL_ISO, B_RADIAL_ADI, B_AZIM_ISO, B_AZIM_ADI,
    B_VELOCITIES, B_DAMPING, B_DERIVED, B_SRC_FUSED, B_AV_FUSED, B_VISC_FUSED, B_FINAL_DERIVED,
    B_CFL_RINGS, B_CFL_FINAL_RINGS, B_KICK_WALK, B_VISC_WALK, B_GHOST_MEAN, B_VEL_FINAL
};
struct BatchStageHost {
    int kind, arg;
    std::vector<BatchEntry> entries;
};
struct BatchStageDev {
    int kind, arg;
    const BatchEntry *map;
    unsigned n;
};
// the stages of one sim::run iteration for a set of disks (parameters as the
// kernels see them, radial_tile resolved); stages no disk takes part in are left out
std::vector<BatchStageHost> plan_batch_step(const std::vector<KParams> &disks);
void launch_batch_stage(const KParams *Ps, const BatchStageDev &s, int split, int radial_waves,
                        hipStream_t st);

}  // namespace fcpt
