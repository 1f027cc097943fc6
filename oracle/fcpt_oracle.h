/*
 * CPU oracle for the FargoCPT hydro step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C (C11 + OpenMP) restatement of the reference's FP64 loops, kept
 * loop-for-loop and operation-for-operation equal to them so that, built
 * without fast-math and without FMA contraction, it reproduces the reference
 * bit for bit.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it, and only as the checker or the timed CPU
 * baseline; the GPU library never calls it.
 *
 * The API mirrors include/fcpt.h (same descriptor, field ids and call order),
 * with host arrays instead of device buffers.
 */
#ifndef FCPT_ORACLE_H
#define FCPT_ORACLE_H

#include "../include/fcpt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fo_disk fo_disk;

fo_disk *fo_create(const fcpt_disk_desc *desc);
void fo_destroy(fo_disk *d);
void fo_set_threads(int n);
int fo_threads(void);
/* direct access to a field array in the reference layout */
double *fo_field(fo_disk *d, int field);
int fo_field_rows(const fo_disk *d, int field);
void fo_set_state(fo_disk *d, const double *sigma, const double *vrad,
                  const double *vazi, const double *energy);
void fo_set_reference(fo_disk *d, const double *sigma0, const double *vrad0,
                      const double *vazi0, const double *energy0);
void fo_set_time(fo_disk *d, double time, double last_dt, int n_monitor);
void fo_get_time(const fo_disk *d, double *time, double *last_dt, int *n_monitor);
void fo_set_step_params(fo_disk *d, const fcpt_step_params *p);
void fo_set_midstep_params(fo_disk *d, const fcpt_step_params *p);
void fo_init_profiles(fo_disk *d, const fcpt_ic_desc *ic);
/* geometry: name is one of Radii Rinf Rsup Rmed Surf InvSurf InvRmed InvRinf
   InvDiffRmed InvDiffRsup InvDiffRsupRb TwoDiffRaSq; returns NULL if unknown */
const double *fo_geom(const fo_disk *d, const char *name);
int fo_damping_ids(const fo_disk *d, int *out4);

double fo_condition_cfl(fo_disk *d);
double fo_calculate_timestep(fo_disk *d);
void fo_nbody_potential(fo_disk *d);
void fo_update_with_sourceterms(fo_disk *d, double dt);
void fo_artificial_viscosity(fo_disk *d, double dt);
void fo_recalculate_viscosity(fo_disk *d);
void fo_viscous_stress_tensor(fo_disk *d);
void fo_update_velocities_with_viscosity(fo_disk *d, double dt);
void fo_substep3(fo_disk *d, double time, double dt);
void fo_apply_boundary_condition(fo_disk *d, double dt, int final_call);
void fo_transport(fo_disk *d, double dt);
void fo_recalculate_derived(fo_disk *d);
const int32_t *fo_nshift(const fo_disk *d);
void fo_step_euler(fo_disk *d, double dt);
void fo_step_leapfrog(fo_disk *d, double dt);
void fo_compute_pressure(fo_disk *d);
double fo_run_step(fo_disk *d);
void fo_totals(fo_disk *d, double radius, double *mass, double *angmom);
void fo_disk_on_planet_accel(fo_disk *d, const fcpt_force_body *bodies, int n_bodies, int flags,
                             double *accel_xy);

#ifdef __cplusplus
}
#endif
#endif
