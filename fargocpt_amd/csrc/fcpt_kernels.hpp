// Launchers of the gfx950 hydro-step kernels (fcpt_kernels.hip).  Every
// launcher enqueues on the given stream and reads step-varying scalars (dt,
// time, planets) from device memory, so a sequence of launchers can be
// captured into a HIP graph and replayed.
#pragma once

#include "fcpt_device.hpp"

namespace fcpt {

void launch_ring_mean(const KParams &P, hipStream_t st, int mode);
void launch_cfl(const KParams &P, hipStream_t st, int mode);
void launch_set_dt(const KParams &P, hipStream_t st, double dt);
void launch_set_clock(const KParams &P, hipStream_t st, double t);
void launch_set_time(const KParams &P, hipStream_t st, double t, double last_dt, int n_monitor);
void launch_set_params(const KParams &P, hipStream_t st, const fcpt_step_params &sp);
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
void launch_radial_sweep(const KParams &P, hipStream_t st);
void launch_azimuthal_sweep(const KParams &P, hipStream_t st);
void launch_cfl_cells(const KParams &P, hipStream_t st);
void launch_copy4(const KParams &P, hipStream_t st);
void launch_derived(const KParams &P, hipStream_t st);
void launch_iso_profile(const KParams &P, hipStream_t st);
void launch_totals(const KParams &P, hipStream_t st, double radius, double *ring_scratch,
                   double *out2);
// split storage: deviations of `field` (rows x nphi floats) re-expressed from
// background oldbg to newbg (both device arrays)
void launch_rebase(const KParams &P, hipStream_t st, void *field, const double *oldbg,
                   const double *newbg, int rows);
// bytes of one stored value of a 2-D field (8 in FP64 storage, 4 in split storage)
size_t element_size(const KParams &P);
// rings per thread of the radial sweep (requested > 0 overrides the size-based choice)
int radial_tile_for(int nr, int ns, int requested);

}  // namespace fcpt
