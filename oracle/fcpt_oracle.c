/*
 * CPU oracle for the FargoCPT hydro step -- TEST INFRASTRUCTURE ONLY
 * (see fcpt_oracle.h).  Every function names the reference function it
 * restates (file:line in the reference tree).  Operation order inside each
 * expression follows the reference's C++ expression exactly; build with
 * -fno-fast-math -ffp-contract=off (oracle/Makefile) to keep it bit-exact.
 */
#include "fcpt_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

struct fo_disk {
    fcpt_disk_desc p;
    int nr, ns;
    double dphi, invdphi, rmin, rmax;
    /* radial arrays (init.cpp:188-225) */
    double *Radii, *Rinf, *Rsup, *Rmed, *Surf, *InvSurf, *InvRmed, *InvRinf,
        *InvDiffRmed, *InvDiffRsup, *InvDiffRsupRb, *TwoDiffRaSq, *cosphi, *sinphi;
    double *f[FCPT_NUM_FIELDS];
    /* transport scratch (TransportEuler.cpp:31-50) */
    double *mom[4], *vres, *work, *qrstar, *densstar, *densint, *dq, *tmp;
    double *vmean;
    int32_t *nshift;
    int damp_ids[4];
    double time, last_dt;
    int n_monitor;
    fcpt_step_params sp;
    fcpt_step_params sp_mid; /* LeapFrog second kick (has_mid) */
    int has_mid;
};

static int rows_of(int field, int nr)
{
    return (field == FCPT_VRAD || field == FCPT_VRAD0 || field == FCPT_TAU_R_PHI) ? nr + 1 : nr;
}

int fo_field_rows(const fo_disk *d, int field) { return rows_of(field, d->nr); }
double *fo_field(fo_disk *d, int field) { return d->f[field]; }
const int32_t *fo_nshift(const fo_disk *d) { return d->nshift; }

#define IDX(i, j) ((size_t)(i) * (size_t)ns + (size_t)(j))

/* std::min / std::max semantics (the second argument wins only if strictly
   smaller / larger), which fixes the sign of zero results like the reference */
static inline double std_min(double a, double b) { return (b < a) ? b : a; }
static inline double std_max(double a, double b) { return (a < b) ? b : a; }

static double omega_kepler(const fo_disk *d, double r)
{
    /* calculate_omega_kepler, Theo.cpp:246-249 */
    return sqrt(d->p.G * d->p.hydro_center_mass / (r * r * r));
}

/* get_rmed_id / get_rinf_id for the logarithmic mesh with the NDEBUG clamp
   (find_cell_id.cpp:218-224,259-264,107-215) */
static int rmed_id_log(const fo_disk *d, double r)
{
    const double g = pow(d->rmax / d->rmin, 1.0 / ((double)d->nr - 2.0));
    const double opt = 3.0 / 2.0 / d->rmin * (1 - pow(g, 2.0)) / (1 - pow(g, 3.0));
    const double inv_log = 1.0 / log(g);
    const double did = log(r * opt) * inv_log;
    int id = (int)floor(did) + 1;
    if (id < 0) id = 0;
    if (id > d->nr - 1) id = d->nr - 1;
    return id;
}

static int rinf_id_log(const fo_disk *d, double r)
{
    const double g = pow(d->rmax / d->rmin, 1.0 / ((double)d->nr - 2.0));
    const double inv_log = 1.0 / log(g);
    const double did = log(r / d->rmin) * inv_log;
    int id = (int)floor(did) + 1;
    if (id < 0) id = 0;
    if (id > d->nr - 1) id = d->nr - 1;
    return id;
}

static int clamp_id(int id, int max)
{
    if (id < 0) return 0;
    if (id > max) return max;
    return id;
}

fo_disk *fo_create(const fcpt_disk_desc *desc)
{
    fo_disk *d = (fo_disk *)calloc(1, sizeof(fo_disk));
    d->p = *desc;
    d->p.radii = NULL;
    const int nr = desc->nr, ns = desc->nphi;
    d->nr = nr;
    d->ns = ns;
    /* Interpret.cpp:230-231 */
    d->dphi = 2.0 * M_PI / (double)ns;
    d->invdphi = (double)ns / (2.0 * M_PI);
    d->rmin = desc->rmin;
    d->rmax = desc->rmax;
    const int n1 = nr + 2;
    d->Radii = calloc(n1, sizeof(double));
    d->Rinf = calloc(n1, sizeof(double));
    d->Rsup = calloc(n1, sizeof(double));
    d->Rmed = calloc(n1, sizeof(double));
    d->Surf = calloc(n1, sizeof(double));
    d->InvSurf = calloc(n1, sizeof(double));
    d->InvRmed = calloc(n1, sizeof(double));
    d->InvRinf = calloc(n1, sizeof(double));
    d->InvDiffRmed = calloc(n1, sizeof(double));
    d->InvDiffRsup = calloc(n1, sizeof(double));
    d->InvDiffRsupRb = calloc(n1, sizeof(double));
    d->TwoDiffRaSq = calloc(n1, sizeof(double));
    d->cosphi = calloc(ns, sizeof(double));
    d->sinphi = calloc(ns, sizeof(double));
    memcpy(d->Radii, desc->radii, sizeof(double) * (nr + 2));
    /* init_radialarrays, init.cpp:188-225 (single rank: IMIN = 0) */
    for (int i = 0; i < nr + 1; ++i) {
        d->Rinf[i] = d->Radii[i];
        d->Rsup[i] = d->Radii[i + 1];
        d->Rmed[i] = 2.0 / 3.0 * (pow(d->Rsup[i], 3) - pow(d->Rinf[i], 3));
        d->Rmed[i] = d->Rmed[i] / (pow(d->Rsup[i], 2) - pow(d->Rinf[i], 2));
        d->Surf[i] = M_PI * (pow(d->Rsup[i], 2) - pow(d->Rinf[i], 2)) / (double)ns;
        d->InvRmed[i] = 1.0 / d->Rmed[i];
        d->InvSurf[i] = 1.0 / d->Surf[i];
        d->InvDiffRsup[i] = 1.0 / (d->Rsup[i] - d->Rinf[i]);
        d->InvDiffRsupRb[i] = 1.0 / ((d->Rsup[i] - d->Rinf[i]) * d->Rmed[i]);
        d->TwoDiffRaSq[i] = 2.0 / (d->Rsup[i] * d->Rsup[i] - d->Rinf[i] * d->Rinf[i]);
        d->InvRinf[i] = 1.0 / d->Rinf[i];
    }
    for (int i = 1; i < nr + 1; ++i)
        d->InvDiffRmed[i] = 1.0 / (d->Rmed[i] - d->Rmed[i - 1]);
    /* cell centres (SideEuler.cpp:57-66) use cos/sin of dphi*j */
    for (int j = 0; j < ns; ++j) {
        d->cosphi[j] = cos(d->dphi * (double)j);
        d->sinphi[j] = sin(d->dphi * (double)j);
    }
    for (int k = 0; k < FCPT_NUM_FIELDS; ++k)
        d->f[k] = calloc((size_t)rows_of(k, nr) * ns, sizeof(double));
    for (int k = 0; k < 4; ++k) d->mom[k] = calloc((size_t)nr * ns, sizeof(double));
    d->vres = calloc((size_t)nr * ns, sizeof(double));
    d->work = calloc((size_t)nr * ns, sizeof(double));
    d->qrstar = calloc((size_t)(nr + 1) * ns, sizeof(double));
    d->densstar = calloc((size_t)(nr + 1) * ns, sizeof(double));
    d->densint = calloc((size_t)nr * ns, sizeof(double));
    d->dq = calloc((size_t)nr * ns, sizeof(double));
    d->tmp = calloc((size_t)nr * ns, sizeof(double));
    d->vmean = calloc(nr + 1, sizeof(double));
    d->nshift = calloc(nr, sizeof(int32_t));
    /* damping ring limits (damping.cpp:321-327, 377-383) */
    int ids[4] = {desc->damp_inner_scalar, desc->damp_inner_vector,
                  desc->damp_outer_scalar, desc->damp_outer_vector};
    if (ids[0] < 0) ids[0] = clamp_id(rmed_id_log(d, d->rmin * desc->damping_inner_limit), nr - 1);
    if (ids[1] < 0) ids[1] = clamp_id(rinf_id_log(d, d->rmin * desc->damping_inner_limit), nr);
    if (ids[2] < 0)
        ids[2] = clamp_id(rmed_id_log(d, d->rmax * desc->damping_outer_limit) + 1, nr - 1);
    if (ids[3] < 0)
        ids[3] = clamp_id(rinf_id_log(d, d->rmax * desc->damping_outer_limit) + 1, nr);
    memcpy(d->damp_ids, ids, sizeof(ids));
    return d;
}

/* OpenMP threads of the oracle's loops (the timed CPU baseline states them) */
void fo_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int fo_threads(void) { return omp_get_max_threads(); }

void fo_destroy(fo_disk *d)
{
    if (!d) return;
    double *arrs[] = {d->Radii, d->Rinf, d->Rsup, d->Rmed, d->Surf, d->InvSurf, d->InvRmed,
                      d->InvRinf, d->InvDiffRmed, d->InvDiffRsup, d->InvDiffRsupRb,
                      d->TwoDiffRaSq, d->cosphi, d->sinphi, d->vres, d->work, d->qrstar,
                      d->densstar, d->densint, d->dq, d->tmp, d->vmean};
    for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); ++k) free(arrs[k]);
    for (int k = 0; k < FCPT_NUM_FIELDS; ++k) free(d->f[k]);
    for (int k = 0; k < 4; ++k) free(d->mom[k]);
    free(d->nshift);
    free(d);
}

int fo_damping_ids(const fo_disk *d, int *out4)
{
    memcpy(out4, d->damp_ids, sizeof(d->damp_ids));
    return 0;
}

const double *fo_geom(const fo_disk *d, const char *name)
{
    struct { const char *n; const double *a; } t[] = {
        {"Radii", d->Radii}, {"Rinf", d->Rinf}, {"Rsup", d->Rsup}, {"Rmed", d->Rmed},
        {"Surf", d->Surf}, {"InvSurf", d->InvSurf}, {"InvRmed", d->InvRmed},
        {"InvRinf", d->InvRinf}, {"InvDiffRmed", d->InvDiffRmed},
        {"InvDiffRsup", d->InvDiffRsup}, {"InvDiffRsupRb", d->InvDiffRsupRb},
        {"TwoDiffRaSq", d->TwoDiffRaSq}};
    for (size_t k = 0; k < sizeof(t) / sizeof(t[0]); ++k)
        if (strcmp(t[k].n, name) == 0) return t[k].a;
    return NULL;
}

void fo_set_time(fo_disk *d, double time, double last_dt, int n_monitor)
{
    d->time = time;
    d->last_dt = last_dt;
    d->n_monitor = n_monitor;
}

void fo_get_time(const fo_disk *d, double *time, double *last_dt, int *n_monitor)
{
    *time = d->time;
    *last_dt = d->last_dt;
    *n_monitor = d->n_monitor;
}

void fo_set_step_params(fo_disk *d, const fcpt_step_params *p)
{
    d->sp = *p;
    d->has_mid = 0;
}

void fo_set_midstep_params(fo_disk *d, const fcpt_step_params *p)
{
    d->sp_mid = *p;
    d->has_mid = 1;
}

/* ---------------------------------------------------------------------- */
/* derived fields                                                          */

/* compute_temperature, SourceEuler.cpp:1378-1408 */
static void compute_temperature(fo_disk *d)
{
    const int nr = d->nr, ns = d->ns;
    double *T = d->f[FCPT_TEMPERATURE], *S = d->f[FCPT_SIGMA], *E = d->f[FCPT_ENERGY],
           *P = d->f[FCPT_PRESSURE];
    const double Rgas = d->p.R;
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            if (d->p.adiabatic) {
                const double c_v_inv = d->p.mu / Rgas * (d->p.gamma - 1.0);
                T[IDX(i, j)] = c_v_inv * E[IDX(i, j)] / S[IDX(i, j)];
            } else {
                T[IDX(i, j)] = d->p.mu / Rgas * P[IDX(i, j)] / S[IDX(i, j)];
            }
        }
}

/* compute_sound_speed_normal, SourceEuler.cpp:957-995 */
static void compute_sound_speed(fo_disk *d)
{
    const int nr = d->nr, ns = d->ns;
    double *cs = d->f[FCPT_SOUNDSPEED], *S = d->f[FCPT_SIGMA], *E = d->f[FCPT_ENERGY];
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            if (d->p.adiabatic) {
                const double g = d->p.gamma;
                cs[IDX(i, j)] = sqrt(g * (g - 1.0) * E[IDX(i, j)] / S[IDX(i, j)]);
            } else {
                const double vK = sqrt(d->p.G * d->p.hydro_center_mass / d->Rmed[i]);
                const double h = d->p.aspect_ratio * pow(d->Rmed[i], d->p.flaring_index);
                cs[IDX(i, j)] = h * vK;
            }
        }
}

/* compute_scale_height_old, SourceEuler.cpp:1121-1154 */
static void compute_scale_height(fo_disk *d)
{
    const int nr = d->nr, ns = d->ns;
    double *H = d->f[FCPT_SCALE_HEIGHT], *cs = d->f[FCPT_SOUNDSPEED];
#pragma omp parallel for
    for (int i = 0; i < nr; ++i) {
        const double inv_omega_kepler = 1.0 / omega_kepler(d, d->Rmed[i]);
        for (int j = 0; j < ns; ++j) {
            if (d->p.adiabatic)
                H[IDX(i, j)] = cs[IDX(i, j)] / (sqrt(d->p.gamma)) * inv_omega_kepler;
            else
                H[IDX(i, j)] = cs[IDX(i, j)] * inv_omega_kepler;
        }
    }
}

/* compute_pressure, SourceEuler.cpp:1345-1376 */
static void compute_pressure(fo_disk *d)
{
    const int nr = d->nr, ns = d->ns;
    double *P = d->f[FCPT_PRESSURE], *S = d->f[FCPT_SIGMA], *E = d->f[FCPT_ENERGY],
           *cs = d->f[FCPT_SOUNDSPEED];
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            if (d->p.adiabatic)
                P[IDX(i, j)] = (d->p.gamma - 1.0) * E[IDX(i, j)];
            else
                P[IDX(i, j)] = S[IDX(i, j)] * pow(cs[IDX(i, j)], 2);
        }
}

/* viscosity::update_viscosity (CONST_ALPHA), viscosity.cpp:98-137 */
static void update_viscosity(fo_disk *d)
{
    const int nr = d->nr, ns = d->ns;
    double *nu = d->f[FCPT_VISCOSITY], *cs = d->f[FCPT_SOUNDSPEED], *H = d->f[FCPT_SCALE_HEIGHT];
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            const double alpha = d->p.alpha;
            nu[IDX(i, j)] = alpha * H[IDX(i, j)] * cs[IDX(i, j)];
        }
}

/* assure_temperature_range, SourceEuler.cpp:136-202 */
static void assure_temperature_range(fo_disk *d)
{
    const int nr = d->nr, ns = d->ns;
    double *E = d->f[FCPT_ENERGY], *S = d->f[FCPT_SIGMA];
    const double Tmin = d->p.t_min, Tmax = d->p.t_max, mu = d->p.mu, g = d->p.gamma,
                 R = d->p.R;
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            const double minimum_energy = Tmin * S[IDX(i, j)] / mu * R / (g - 1.0);
            const double maximum_energy = Tmax * S[IDX(i, j)] / mu * R / (g - 1.0);
            if (!(E[IDX(i, j)] > minimum_energy))
                E[IDX(i, j)] = Tmin * S[IDX(i, j)] / mu * R / (g - 1.0);
            if (!(E[IDX(i, j)] < maximum_energy))
                E[IDX(i, j)] = Tmax * S[IDX(i, j)] / mu * R / (g - 1.0);
        }
}

/* calculate_qminus / calculate_qplus / the CFL heating-cooling division
   (SourceEuler.cpp:632-690, 496-536, 1410-1450) */
static void thermal_relaxation(fo_disk *d, double current_time);
static void viscous_heating(fo_disk *d);

static void heating_cooling_for_cfl(fo_disk *d)
{
    /* compute_heating_cooling_for_CFL, SourceEuler.cpp:1410-1450 */
    const int nr = d->nr, ns = d->ns;
    if (!d->p.adiabatic) return;
    update_viscosity(d);
    fo_viscous_stress_tensor(d);
    memset(d->f[FCPT_QMINUS], 0, sizeof(double) * nr * ns);
    if (d->p.cooling_beta_enabled) thermal_relaxation(d, d->time);
    memset(d->f[FCPT_QPLUS], 0, sizeof(double) * nr * ns);
    if (d->p.heating_viscous) viscous_heating(d);
    double *Qp = d->f[FCPT_QPLUS], *Qm = d->f[FCPT_QMINUS], *H = d->f[FCPT_SCALE_HEIGHT],
           *S = d->f[FCPT_SIGMA], *E = d->f[FCPT_ENERGY];
#pragma omp parallel for
    for (int i = 1; i < nr - 1; ++i)
        for (int j = 0; j < ns; ++j) {
            const double inv_pow4 =
                pow(d->p.mu * (d->p.gamma - 1.0) / (d->p.R * S[IDX(i, j)]), 4);
            double alpha = 1.0 + 2.0 * H[IDX(i, j)] * 4.0 * d->p.sigma_sb / d->p.c_light *
                                     inv_pow4 * pow(E[IDX(i, j)], 3);
            Qp[IDX(i, j)] /= alpha;
            Qm[IDX(i, j)] /= alpha;
        }
}

void fo_set_state(fo_disk *d, const double *sigma, const double *vrad, const double *vazi,
                  const double *energy)
{
    const size_t n = (size_t)d->nr * d->ns;
    memcpy(d->f[FCPT_SIGMA], sigma, n * sizeof(double));
    memcpy(d->f[FCPT_VRAD], vrad, (n + d->ns) * sizeof(double));
    memcpy(d->f[FCPT_VAZI], vazi, n * sizeof(double));
    if (energy) memcpy(d->f[FCPT_ENERGY], energy, n * sizeof(double));
    /* init_euler recalculation, SourceEuler.cpp:264-284 */
    if (!d->p.adiabatic) {
        compute_sound_speed(d);
        compute_pressure(d);
        compute_temperature(d);
        compute_scale_height(d);
    } else {
        compute_temperature(d);
        compute_sound_speed(d);
        compute_scale_height(d);
        compute_pressure(d);
    }
    update_viscosity(d);
    heating_cooling_for_cfl(d);
}

/* ---- initial conditions (init_physics, init.cpp:255-343, profile path) ---- */

/* cutoff_outer / cutoff_inner, util.cpp:69-93 */
static double ic_cutoff_outer(double point, double width, double x)
{
    return 1.0 / (1.0 + exp((x - point) / width));
}
static double ic_cutoff_inner(double point, double width, double x)
{
    return 1.0 / (1.0 + exp((point - x) / width));
}

/* support_azi_smoothing_derivative, support_azi_pressure and
   initial_locally_isothermal_smoothed_v_az, Theo.cpp:131-181 */
static double ic_vaz(const fo_disk *d, const fcpt_ic_desc *ic, double R)
{
    const double h0 = d->p.aspect_ratio, F = d->p.flaring_index, S = ic->sigma_slope;
    const double eps = d->p.thickness_smoothing;
    const double h = h0 * pow(R, F);
    const double smoothing_derivative_2 =
        (1.0 + (F + 1.0) * pow(h * eps, 2)) / pow(sqrt(1 + pow(h * eps, 2)), 3);
    const double pressure_support_2 = (2.0 * F - 1.0 - S) * pow(h, 2);
    const double support = smoothing_derivative_2 + pressure_support_2;
    const double vk_2 = d->p.G * d->p.hydro_center_mass / R;
    return sqrt(vk_2 * support);
}

/* get_sigma, viscous_radial_speed.cpp:95-118 */
static double ic_sigma(const fo_disk *d, const fcpt_ic_desc *ic, double R)
{
    double density = d->p.sigma0 * pow(R, -ic->sigma_slope);
    const double density_floor = d->p.sigma_floor * d->p.sigma0;
    if (ic->cutoff_outer) density *= ic_cutoff_outer(ic->cutoff_outer_point, ic->cutoff_outer_width, R);
    if (ic->cutoff_inner) density *= ic_cutoff_inner(ic->cutoff_inner_point, ic->cutoff_inner_width, R);
    return std_max(density, density_floor);
}

/* get_nu2, viscous_radial_speed.cpp:38-93 */
static double ic_nu(const fo_disk *d, const fcpt_ic_desc *ic, double R, double Sigma)
{
    const double M = d->p.hydro_center_mass;
    const double v_k = sqrt(d->p.G * M / R);
    const double h = d->p.aspect_ratio * pow(R, d->p.flaring_index);
    double cutoff = 1.0;
    if (ic->cutoff_outer) cutoff *= ic_cutoff_outer(ic->cutoff_outer_point, ic->cutoff_outer_width, R);
    if (ic->cutoff_inner) cutoff *= ic_cutoff_inner(ic->cutoff_inner_point, ic->cutoff_inner_width, R);
    double cs_adb, H;
    if (d->p.adiabatic) {
        const double gamma = d->p.gamma;
        double energy = cutoff * 1.0 / (gamma - 1.0) * Sigma * pow(h * v_k, 2);
        const double energy_floor = d->p.t_min * Sigma / d->p.mu * d->p.R / (gamma - 1.0);
        const double energy_ceil = d->p.t_max * Sigma / d->p.mu * d->p.R / (gamma - 1.0);
        energy = std_max(energy, energy_floor);
        energy = std_min(energy, energy_ceil);
        cs_adb = sqrt(gamma * (gamma - 1.0) * energy / Sigma);
        const double cs_iso = sqrt((gamma - 1.0) * energy / Sigma);
        const double omega_k = v_k / R;
        H = cs_iso / omega_k;
    } else {
        cs_adb = h * v_k;
        H = h * R;
    }
    return d->p.alpha * cs_adb * H;
}

/* the functions viscous_speed differentiates (viscous_radial_speed.cpp:140-172) */
enum { IC_W, IC_R2W, IC_NU_S_R3_DWDR };
static double ic_eval(const fo_disk *d, const fcpt_ic_desc *ic, int which, double r);

/* derive, viscous_radial_speed.cpp:120-138: five-point first derivative */
static double ic_derive(const fo_disk *d, const fcpt_ic_desc *ic, int which, double r)
{
    const double x = r;
    const double h = 8.0e-4 * x;
    const double f1 = -1.0 * ic_eval(d, ic, which, x + 2.0 * h);
    const double f2 = 8.0 * ic_eval(d, ic, which, x + h);
    const double f3 = -8.0 * ic_eval(d, ic, which, x - h);
    const double f4 = 1.0 * ic_eval(d, ic, which, x - 2.0 * h);
    return (f1 + f2 + f3 + f4) / (12.0 * h);
}

static double ic_eval(const fo_disk *d, const fcpt_ic_desc *ic, int which, double r)
{
    if (which == IC_W) return ic_vaz(d, ic, r) / r;
    if (which == IC_R2W) return pow(r, 2) * (ic_vaz(d, ic, r) / r);
    const double dw_dr = ic_derive(d, ic, IC_W, r);
    const double Sigma = ic_sigma(d, ic, r);
    const double nu = ic_nu(d, ic, r, Sigma);
    return nu * Sigma * pow(r, 3) * dw_dr;
}

/* get_vr_with_numerical_viscous_speed, viscous_radial_speed.cpp:174-195 */
static double ic_vr_visc(const fo_disk *d, const fcpt_ic_desc *ic, double r)
{
    const double num = 1.0 / r * ic_derive(d, ic, IC_NU_S_R3_DWDR, r);
    const double Sigma = ic_sigma(d, ic, r);
    const double den = Sigma * ic_derive(d, ic, IC_R2W, r);
    return num / den;
}

void fo_init_profiles(fo_disk *d, const fcpt_ic_desc *ic)
{
    const int nr = d->nr, ns = d->ns;
    double *S = d->f[FCPT_SIGMA], *E = d->f[FCPT_ENERGY], *vr = d->f[FCPT_VRAD],
           *vaz = d->f[FCPT_VAZI];
    const double floor_value = d->p.sigma_floor * d->p.sigma0;
    /* init_gas_density, init.cpp:937-1148 (profile, cutoffs) */
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            const double density = d->p.sigma0 * pow(d->Rmed[i], -ic->sigma_slope);
            S[IDX(i, j)] = std_max(density, floor_value);
        }
    if (ic->cutoff_outer)
        for (int i = 0; i < nr; ++i)
            for (int j = 0; j < ns; ++j) {
                const double damped = S[IDX(i, j)] * ic_cutoff_outer(ic->cutoff_outer_point,
                                                                    ic->cutoff_outer_width, d->Rmed[i]);
                S[IDX(i, j)] = std_max(damped, floor_value);
            }
    if (ic->cutoff_inner)
        for (int i = 0; i < nr; ++i)
            for (int j = 0; j < ns; ++j) {
                const double damped = S[IDX(i, j)] * ic_cutoff_inner(ic->cutoff_inner_point,
                                                                    ic->cutoff_inner_width, d->Rmed[i]);
                S[IDX(i, j)] = std_max(damped, floor_value);
            }
    /* init_gas_energy, init.cpp:1257-1300 (profile) with initial_energy, Theo.cpp:86-99 */
    if (d->p.adiabatic) {
        const double h0 = d->p.aspect_ratio, F = d->p.flaring_index, Sl = ic->sigma_slope;
        for (int i = 0; i < nr; ++i)
            for (int j = 0; j < ns; ++j) {
                const double energy = 1.0 / (d->p.gamma - 1.0) * d->p.sigma0 * pow(h0, 2) *
                                      pow(d->Rmed[i], -Sl - 1.0 + 2.0 * F) * d->p.G *
                                      d->p.hydro_center_mass;
                const double energy_floor =
                    d->p.t_min * S[IDX(i, j)] / d->p.mu * d->p.R / (d->p.gamma - 1.0);
                E[IDX(i, j)] = std_max(energy, energy_floor);
            }
    }
    /* renormalize_sigma_and_report without SigmaAdjust: SigmaMed, SigmaInf
       (compute_azi_avg_Sigma, Theo.cpp:30-48) */
    double *smed = (double *)calloc((size_t)nr + 1, sizeof(double));
    double *sinf = (double *)calloc((size_t)nr + 1, sizeof(double));
    for (int i = 0; i < nr; ++i) {
        double mean = 0.0;
        for (int j = 0; j < ns; ++j) mean += S[IDX(i, j)];
        smed[i] = mean / (double)ns;
    }
    sinf[0] = smed[0];
    for (int i = 1; i < nr; ++i) {
        const double sigm = smed[i - 1], sigp = smed[i];
        const double dr = d->Rmed[i] - d->Rmed[i - 1];
        sinf[i] = (sigm * (d->Rmed[i] - d->Rinf[i]) + sigp * (d->Rinf[i] - d->Rmed[i - 1])) / dr;
    }
    /* init_euler (SourceEuler.cpp:251-284) while the velocities are still zero */
    memset(vr, 0, sizeof(double) * (size_t)(nr + 1) * ns);
    memset(vaz, 0, sizeof(double) * (size_t)nr * ns);
    if (!d->p.adiabatic) {
        compute_sound_speed(d);
        compute_pressure(d);
        compute_temperature(d);
        compute_scale_height(d);
    } else {
        compute_temperature(d);
        compute_sound_speed(d);
        compute_scale_height(d);
        compute_pressure(d);
    }
    update_viscosity(d);
    heating_cooling_for_cfl(d);
    /* init_gas_velocities, init.cpp:1725-1770 (no self-gravity, no centrifugal
       balance) */
    for (int i = 0; i < nr; ++i) {
        const double r = d->Rmed[i], ri = d->Rinf[i];
        for (int j = 0; j < ns; ++j) {
            vaz[IDX(i, j)] = ic_vaz(d, ic, r);
            vaz[IDX(i, j)] -= d->p.omega_frame * r;
            vr[IDX(i, j)] = ic->imposed_disk_drift * d->p.sigma0 / sinf[i] / ri;
            if (!ic->vradial_zero) vr[IDX(i, j)] += ic_vr_visc(d, ic, ri);
            else vr[IDX(i, j)] = 0.0;
        }
    }
    free(smed);
    free(sinf);
    /* copy_initial_values, the boundary call, copy_initial_values (init.cpp:337-341) */
    const size_t n = (size_t)nr * ns;
    memcpy(d->f[FCPT_SIGMA0], S, n * sizeof(double));
    memcpy(d->f[FCPT_VRAD0], vr, (n + ns) * sizeof(double));
    memcpy(d->f[FCPT_VAZI0], vaz, n * sizeof(double));
    memcpy(d->f[FCPT_ENERGY0], E, n * sizeof(double));
    fo_apply_boundary_condition(d, 0.0, 0);
    memcpy(d->f[FCPT_SIGMA0], S, n * sizeof(double));
    memcpy(d->f[FCPT_VRAD0], vr, (n + ns) * sizeof(double));
    memcpy(d->f[FCPT_VAZI0], vaz, n * sizeof(double));
    memcpy(d->f[FCPT_ENERGY0], E, n * sizeof(double));
}

void fo_set_reference(fo_disk *d, const double *sigma0, const double *vrad0,
                      const double *vazi0, const double *energy0)
{
    const size_t n = (size_t)d->nr * d->ns;
    memcpy(d->f[FCPT_SIGMA0], sigma0, n * sizeof(double));
    memcpy(d->f[FCPT_VRAD0], vrad0, (n + d->ns) * sizeof(double));
    memcpy(d->f[FCPT_VAZI0], vazi0, n * sizeof(double));
    memcpy(d->f[FCPT_ENERGY0], energy0, n * sizeof(double));
}

/* ---------------------------------------------------------------------- */
/* CFL: cfl::condition_cfl, cfl.cpp:185-382 (single rank, Euler integrator) */
double fo_condition_cfl(fo_disk *d)
{
    const int nr = d->nr, ns = d->ns;
    const double *vr = d->f[FCPT_VRAD], *vaz = d->f[FCPT_VAZI], *cs = d->f[FCPT_SOUNDSPEED],
                 *nu = d->f[FCPT_VISCOSITY], *Qp = d->f[FCPT_QPLUS], *Qm = d->f[FCPT_QMINUS],
                 *E = d->f[FCPT_ENERGY];
    double *v_mean = d->vmean;
    const double CFL = d->p.cfl, dphi = d->dphi;
#pragma omp parallel for
    for (int i = 0; i < nr; ++i) {
        v_mean[i] = 0.0;
        for (int j = 0; j < ns; ++j) v_mean[i] += vaz[IDX(i, j)];
        v_mean[i] /= (double)ns;
    }
    const double denom0 = fabs(v_mean[0] * d->InvRmed[0] - v_mean[1] * d->InvRmed[1]) + 1.0e-100;
    double dt_core = CFL * dphi / denom0;
    const double fac2 = pow(d->p.artvisc_factor, 2);
    /* leapfrog_cfl_factor, cfl.cpp:254-261 */
    const double lf = d->p.integrator == FCPT_INTEGRATOR_LEAPFROG ? 0.6 : 1.0;
#pragma omp parallel for reduction(min : dt_core)
    for (int i = 1; i < nr - 1; ++i) {
        const double denom =
            fabs(v_mean[i] * d->InvRmed[i] - v_mean[i + 1] * d->InvRmed[i + 1]) + 1.0e-100;
        const double shear_dt = CFL * dphi / denom;
        if (shear_dt < dt_core) dt_core = shear_dt;
        const double dxRadial = d->Rsup[i] - d->Rinf[i];
        const double dxAzimuthal = d->Rmed[i] * dphi;
        const double cell_size = std_min(dxRadial, dxAzimuthal);
        for (int j = 0; j < ns; ++j) {
            const double v_residual = vaz[IDX(i, j)] - v_mean[i];
            const double invdt1 = cs[IDX(i, j)] / cell_size;
            const double invdt2 = vr[IDX(i, j)] / dxRadial;
            const double invdt3 = v_residual / dxAzimuthal;
            const int jn = j == ns - 1 ? 0 : j + 1;
            double invdt4;
            if (d->p.artificial_viscosity == FCPT_ARTVISC_SN) {
                double dvRadial = vr[IDX(i + 1, j)] - vr[IDX(i, j)];
                double dvAzimuthal = vaz[IDX(i, jn)] - vaz[IDX(i, j)];
                if (dvRadial > 0.0) dvRadial = 0.0;
                else dvRadial = -dvRadial;
                if (dvAzimuthal > 0.0) dvAzimuthal = 0.0;
                else dvAzimuthal = -dvAzimuthal;
                const double a = dvRadial / dxRadial, b = dvAzimuthal / dxAzimuthal;
                invdt4 = 4.0 * fac2 * std_max(a, b) * lf;
            } else {
                const double eps_rr = (vr[IDX(i + 1, j)] - vr[IDX(i, j)]) * d->InvDiffRsup[i];
                const double eps_pp =
                    d->InvRmed[i] * ((vaz[IDX(i, jn)] - vaz[IDX(i, j)]) * d->invdphi +
                                     0.5 * (vr[IDX(i + 1, j)] + vr[IDX(i, j)]));
                const double mdiv_V = -std_min(eps_rr + eps_pp, 0.0);
                invdt4 = 4.0 * fac2 * mdiv_V * lf;
            }
            const double invdt5 = 4.0 * nu[IDX(i, j)] / pow(cell_size, 2) * lf;
            double invdt6 = 0.0;
            if (d->p.adiabatic) {
                const double inv_limit = 1.0 / d->p.heating_cooling_cfl_limit;
                invdt6 = inv_limit * fabs((Qp[IDX(i, j)] - Qm[IDX(i, j)]) / E[IDX(i, j)]) * lf;
            }
            const double dt_cell =
                CFL / sqrt(pow(invdt1, 2) + pow(invdt2, 2) + pow(invdt3, 2) + pow(invdt4, 2) +
                           pow(invdt5, 2) + pow(invdt6, 2));
            if (dt_cell < dt_core) dt_core = dt_cell;
        }
    }
    return dt_core;
}

/* sim::CalculateTimeStep, simulation.cpp:100-118 */
double fo_calculate_timestep(fo_disk *d)
{
    const double cfl_dt = fo_condition_cfl(d);
    const double rv = std_min(d->p.cfl_max_var * d->last_dt, cfl_dt);
    d->last_dt = rv;
    return rv;
}

/* CalculateNbodyPotential, Pframeforce.cpp:21-86 (+ compute_smoothing, Force.cpp:145-159) */
/* compute_smoothing, Force.cpp:145-159 (compute_smoothing_scaleheight :123-130,
   compute_smoothing_iso_planet :132-142 with planet.get_r() = hypot(x, y),
   planet.cpp:159, and aspectratio_ref = AspectRatio, Interpret.cpp:191) */
static double compute_smoothing(const fo_disk *d, int k, double H)
{
    const int flags = d->p.potential_smoothing;
    if ((flags & FCPT_SMOOTHING_NO_STAR) && k == 0) return 0;
    if (flags & FCPT_SMOOTHING_PLANETLOC) {
        const double a = hypot(d->sp.planet[k].x, d->sp.planet[k].y);
        const double scale_height = d->p.aspect_ratio * pow(a, 1 + d->p.flaring_index);
        return d->p.thickness_smoothing * scale_height;
    }
    return d->p.thickness_smoothing * H;
}

void fo_nbody_potential(fo_disk *d)
{
    const int nr = d->nr, ns = d->ns;
    double *pot = d->f[FCPT_POTENTIAL];
    const double *H = d->f[FCPT_SCALE_HEIGHT];
    const fcpt_step_params *sp = &d->sp;
    const double G = d->p.G;
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            const double x = d->Rmed[i] * d->cosphi[j];
            const double y = d->Rmed[i] * d->sinphi[j];
            double acc = 0.0;
            for (int k = 0; k < sp->n_planets; ++k) {
                const double smooth = compute_smoothing(d, k, H[IDX(i, j)]);
                const double dx = x - sp->planet[k].x;
                const double dy = y - sp->planet[k].y;
                const double dist_2 = pow(dx, 2) + pow(dy, 2);
                const double d_smoothed = sqrt(dist_2 + pow(smooth, 2));
                double smooth_factor_klahr = 1.0;
                const double r_sm = sp->planet[k].r_cubic;
                if (r_sm > 0.0 && d_smoothed < r_sm) {
                    smooth_factor_klahr = (pow(d_smoothed / r_sm, 4.0) -
                                           2.0 * pow(d_smoothed / r_sm, 3.0) +
                                           2.0 * d_smoothed / r_sm);
                }
                acc += -G * sp->planet[k].mass / d_smoothed * smooth_factor_klahr;
            }
            acc += -sp->indirect_x * x - sp->indirect_y * y;
            pot[IDX(i, j)] = acc;
        }
}

/* update_with_sourceterms, SourceEuler.cpp:325-493 */
void fo_update_with_sourceterms(fo_disk *d, double dt)
{
    const int nr = d->nr, ns = d->ns;
    double *S = d->f[FCPT_SIGMA], *P = d->f[FCPT_PRESSURE], *Phi = d->f[FCPT_POTENTIAL],
           *vaz = d->f[FCPT_VAZI], *vr = d->f[FCPT_VRAD], *E = d->f[FCPT_ENERGY];
    const double OmegaF = d->p.omega_frame;
    /* momentum_update_radial, :325-372 (One_no_ghost_vr = 2 .. MaxMo_no_ghost_vr = nr-1) */
#pragma omp parallel for
    for (int i = 2; i < nr - 1; ++i)
        for (int j = 0; j < ns; ++j) {
            double gradp = 2.0 / (S[IDX(i, j)] + S[IDX(i - 1, j)]);
            gradp *= (P[IDX(i, j)] - P[IDX(i - 1, j)]);
            gradp *= d->InvDiffRmed[i];
            const double gradphi = (Phi[IDX(i, j)] - Phi[IDX(i - 1, j)]) * d->InvDiffRmed[i];
            const int jn = j == ns - 1 ? 0 : j + 1;
            const double vsum = vaz[IDX(i, j)] + vaz[IDX(i, jn)] + vaz[IDX(i - 1, j)] +
                                vaz[IDX(i - 1, jn)];
            const double vt = 0.25 * vsum + d->Rinf[i] * OmegaF;
            const double vt2 = vt * vt;
            const double centrifugal_accel = vt2 * d->InvRinf[i];
            vr[IDX(i, j)] += dt * (-gradp - gradphi + centrifugal_accel);
        }
    /* momentum_update_azimuthal, :375-428 (Zero_no_ghost = 1 .. Max_no_ghost = nr-1) */
#pragma omp parallel for
    for (int i = 1; i < nr - 1; ++i) {
        const double invdxtheta = 2.0 / (d->dphi * (d->Rsup[i] + d->Rinf[i]));
        for (int j = 0; j < ns; ++j) {
            const int jp = j == 0 ? ns - 1 : j - 1;
            const double gradp = 2.0 / (S[IDX(i, j)] + S[IDX(i, jp)]) *
                                 (P[IDX(i, j)] - P[IDX(i, jp)]) * invdxtheta;
            const double gradphi = (Phi[IDX(i, j)] - Phi[IDX(i, jp)]) * invdxtheta;
            vaz[IDX(i, j)] = vaz[IDX(i, j)] + dt * (-gradp - gradphi);
        }
    }
    /* compression_heating, :459-493 */
    if (d->p.adiabatic) {
#pragma omp parallel for
        for (int i = 0; i < nr - 1; ++i)
            for (int j = 0; j < ns; ++j) {
                const int jn = j == ns - 1 ? 0 : j + 1;
                const double DIV_V =
                    (vr[IDX(i + 1, j)] * d->Rinf[i + 1] - vr[IDX(i, j)] * d->Rinf[i]) *
                        d->InvDiffRsupRb[i] +
                    (vaz[IDX(i, jn)] - vaz[IDX(i, j)]) * d->invdphi * d->InvRmed[i];
                const double gamma = d->p.gamma;
                const double energy_old = E[IDX(i, j)];
                const double energy_new = energy_old * exp(-(gamma - 1.0) * dt * DIV_V);
                E[IDX(i, j)] = energy_new;
            }
    }
}

/* art_visc::update_with_artificial_viscosity (+_TW / _SN), artificial_viscosity.cpp:11-250 */
void fo_artificial_viscosity(fo_disk *d, double dt)
{
    const int nr = d->nr, ns = d->ns;
    double *E = d->f[FCPT_ENERGY], *S = d->f[FCPT_SIGMA], *vr = d->f[FCPT_VRAD],
           *vaz = d->f[FCPT_VAZI], *Qr = d->f[FCPT_Q_R], *Qp = d->f[FCPT_Q_PHI];
    const double fac = d->p.artvisc_factor;
    if (d->p.artificial_viscosity == FCPT_ARTVISC_TW) {
        /* :35-140 */
#pragma omp parallel for
        for (int i = 0; i < nr; ++i)
            for (int j = 0; j < ns; ++j) {
                const int jn = j == ns - 1 ? 0 : j + 1;
                const double eps_rr = (vr[IDX(i + 1, j)] - vr[IDX(i, j)]) * d->InvDiffRsup[i];
                const double eps_pp =
                    d->InvRmed[i] * ((vaz[IDX(i, jn)] - vaz[IDX(i, j)]) * d->invdphi +
                                     0.5 * (vr[IDX(i + 1, j)] + vr[IDX(i, j)]));
                const double div_V = std_min(eps_rr + eps_pp, 0.0);
                const double Dr = d->Rinf[i + 1] - d->Rinf[i];
                const double rDphi = d->Rmed[i] * d->dphi;
                double dx_sq;
                if (ns <= 16) dx_sq = pow(std_min(Dr, rDphi), 2);
                else dx_sq = pow(std_max(Dr, rDphi), 2);
                const double l_sq = pow(fac, 2) * dx_sq;
                const double q_rr = l_sq * S[IDX(i, j)] * -div_V * (eps_rr - 1.0 / 3.0 * div_V);
                const double q_pp = l_sq * S[IDX(i, j)] * -div_V * (eps_pp - 1.0 / 3.0 * div_V);
                Qr[IDX(i, j)] = q_rr;
                Qp[IDX(i, j)] = q_pp;
                if (d->p.adiabatic && d->p.artvisc_dissipation) {
                    if (i > 1 && i < nr - 1) {
                        const double Qplus = -l_sq * div_V * S[IDX(i, j)] * 1.0 / 3.0 *
                                             (pow(eps_rr, 2) + pow(eps_pp, 2) +
                                              pow((eps_rr - eps_pp), 2));
                        E[IDX(i, j)] += Qplus * dt;
                    }
                }
            }
#pragma omp parallel for
        for (int i = 1; i < nr - 1; ++i)
            for (int j = 0; j < ns; ++j) {
                const int jp = j == 0 ? ns - 1 : j - 1;
                const double sigma_phi_avg = 0.5 * (S[IDX(i, j)] + S[IDX(i, jp)]);
                const double dVp = 2.0 * dt / ((d->Rsup[i] + d->Rinf[i]) * sigma_phi_avg) *
                                   (Qp[IDX(i, j)] - Qp[IDX(i, jp)]) * d->invdphi;
                vaz[IDX(i, j)] += dVp;
            }
#pragma omp parallel for
        for (int i = 2; i < nr - 1; ++i)
            for (int j = 0; j < ns; ++j) {
                const double sigma_r_avg = 0.5 * (S[IDX(i, j)] + S[IDX(i - 1, j)]);
                const double dVr =
                    d->p.radial_viscosity_factor * dt / sigma_r_avg * 2.0 /
                    (pow(d->Rmed[i], 2) - pow(d->Rmed[i - 1], 2)) *
                    ((Qr[IDX(i, j)] * d->Rmed[i] - Qr[IDX(i - 1, j)] * d->Rmed[i - 1]) -
                     0.5 * (Qp[IDX(i, j)] + Qp[IDX(i - 1, j)]) * (d->Rmed[i] - d->Rmed[i - 1]));
                vr[IDX(i, j)] += dVr;
            }
    } else if (d->p.artificial_viscosity == FCPT_ARTVISC_SN) {
        /* :148-250 */
        const double fac2 = pow(fac, 2);
#pragma omp parallel for
        for (int i = 0; i < nr; ++i)
            for (int j = 0; j < ns; ++j) {
                const double dv_r = vr[IDX(i + 1, j)] - vr[IDX(i, j)];
                if (dv_r < 0.0) Qr[IDX(i, j)] = fac2 * S[IDX(i, j)] * pow(dv_r, 2);
                else Qr[IDX(i, j)] = 0.0;
                const int jn = j == ns - 1 ? 0 : j + 1;
                const double dv_phi = vaz[IDX(i, jn)] - vaz[IDX(i, j)];
                if (dv_phi < 0.0) Qp[IDX(i, j)] = fac2 * S[IDX(i, j)] * pow(dv_phi, 2);
                else Qp[IDX(i, j)] = 0.0;
            }
        if (d->p.adiabatic && d->p.artvisc_dissipation) {
#pragma omp parallel for
            for (int i = 1; i < nr - 1; ++i) {
                const double dxtheta = d->dphi * d->Rmed[i];
                const double invdxtheta = 1.0 / dxtheta;
                for (int j = 0; j < ns; ++j) {
                    const int jn = j == ns - 1 ? 0 : j + 1;
                    const double dv_r = vr[IDX(i + 1, j)] - vr[IDX(i, j)];
                    const double dv_phi = vaz[IDX(i, jn)] - vaz[IDX(i, j)];
                    const double energy_new = E[IDX(i, j)] -
                                              dt * Qr[IDX(i, j)] * dv_r * d->InvDiffRsup[i] -
                                              dt * Qp[IDX(i, j)] * dv_phi * invdxtheta;
                    E[IDX(i, j)] = energy_new;
                }
            }
        }
#pragma omp parallel for
        for (int i = 2; i < nr - 1; ++i)
            for (int j = 0; j < ns; ++j)
                vr[IDX(i, j)] = vr[IDX(i, j)] - dt * 2.0 / (S[IDX(i, j)] + S[IDX(i - 1, j)]) *
                                                    (Qr[IDX(i, j)] - Qr[IDX(i - 1, j)]) *
                                                    d->InvDiffRmed[i];
#pragma omp parallel for
        for (int i = 1; i < nr - 1; ++i) {
            const double dxtheta = d->dphi * d->Rmed[i];
            const double invdxtheta = 1.0 / dxtheta;
            for (int j = 0; j < ns; ++j) {
                const int jp = j == 0 ? ns - 1 : j - 1;
                vaz[IDX(i, j)] = vaz[IDX(i, j)] - dt * 2.0 / (S[IDX(i, j)] + S[IDX(i, jp)]) *
                                                      (Qp[IDX(i, j)] - Qp[IDX(i, jp)]) *
                                                      invdxtheta;
            }
        }
    }
    if (d->p.adiabatic && d->p.artvisc_dissipation) assure_temperature_range(d);
}

/* recalculate_viscosity, SourceEuler.cpp:205-223 (AspectRatioMode 0, constant gamma) */
void fo_recalculate_viscosity(fo_disk *d)
{
    if (d->p.adiabatic) {
        compute_sound_speed(d);
        compute_scale_height(d);
    }
    update_viscosity(d);
}

/* viscosity::compute_viscous_stress_tensor, viscosity.cpp:139-254 */
void fo_viscous_stress_tensor(fo_disk *d)
{
    const int nr = d->nr, ns = d->ns;
    double *vr = d->f[FCPT_VRAD], *vaz = d->f[FCPT_VAZI], *nu = d->f[FCPT_VISCOSITY],
           *S = d->f[FCPT_SIGMA], *divv = d->f[FCPT_DIV_V], *trr = d->f[FCPT_TAU_R_R],
           *tpp = d->f[FCPT_TAU_PHI_PHI], *trp = d->f[FCPT_TAU_R_PHI];
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            const int jn = j == ns - 1 ? 0 : j + 1;
            divv[IDX(i, j)] =
                (vr[IDX(i + 1, j)] * d->Rinf[i + 1] - vr[IDX(i, j)] * d->Rinf[i]) *
                    d->InvDiffRsupRb[i] +
                (vaz[IDX(i, jn)] - vaz[IDX(i, j)]) * d->invdphi * d->InvRmed[i];
        }
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            const double drr = (vr[IDX(i + 1, j)] - vr[IDX(i, j)]) * d->InvDiffRsup[i];
            trr[IDX(i, j)] =
                2.0 * nu[IDX(i, j)] * S[IDX(i, j)] * (drr - 1.0 / 3.0 * divv[IDX(i, j)]);
        }
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            const int jn = j == ns - 1 ? 0 : j + 1;
            const double dpp = (vaz[IDX(i, jn)] - vaz[IDX(i, j)]) * d->invdphi * d->InvRmed[i] +
                               0.5 * (vr[IDX(i + 1, j)] + vr[IDX(i, j)]) * d->InvRmed[i];
            const double n = nu[IDX(i, j)];
            const double s = S[IDX(i, j)];
            tpp[IDX(i, j)] = 2.0 * n * s * (dpp - 1.0 / 3.0 * divv[IDX(i, j)]);
        }
#pragma omp parallel for
    for (int i = 1; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            const int jp = j == 0 ? ns - 1 : j - 1;
            const double dvazirdr = (vaz[IDX(i, j)] * d->InvRmed[i] -
                                     vaz[IDX(i - 1, j)] * d->InvRmed[i - 1]) *
                                    d->InvDiffRmed[i];
            const double dvrdphi = (vr[IDX(i, j)] - vr[IDX(i, jp)]) * d->invdphi;
            const double drp = d->Rinf[i] * dvazirdr + dvrdphi * d->InvRinf[i];
            const double n4 = 0.25 * (nu[IDX(i, j)] + nu[IDX(i - 1, j)] + nu[IDX(i, jp)] +
                                      nu[IDX(i - 1, jp)]);
            const double s4 = 0.25 * (S[IDX(i, j)] + S[IDX(i - 1, j)] + S[IDX(i, jp)] +
                                      S[IDX(i - 1, jp)]);
            trp[IDX(i, j)] = n4 * s4 * drp;
        }
}

/* viscosity::update_velocities_with_viscosity, viscosity.cpp:355-426 */
void fo_update_velocities_with_viscosity(fo_disk *d, double dt)
{
    const int nr = d->nr, ns = d->ns;
    double *vr = d->f[FCPT_VRAD], *vaz = d->f[FCPT_VAZI], *S = d->f[FCPT_SIGMA],
           *trr = d->f[FCPT_TAU_R_R], *tpp = d->f[FCPT_TAU_PHI_PHI], *trp = d->f[FCPT_TAU_R_PHI];
    const double invdphi = d->invdphi;
#pragma omp parallel for
    for (int i = 1; i < nr - 1; ++i)
        for (int j = 0; j < ns; ++j) {
            const int jp = j == 0 ? ns - 1 : j - 1;
            const double sigma_avg = 0.5 * (S[IDX(i, j)] + S[IDX(i, jp)]);
            const double dVp =
                dt * d->InvRmed[i] / (sigma_avg) *
                ((2.0 / (pow(d->Rinf[i + 1], 2) - pow(d->Rinf[i], 2))) *
                     (pow(d->Rinf[i + 1], 2) * trp[IDX(i + 1, j)] -
                      pow(d->Rinf[i], 2) * trp[IDX(i, j)]) +
                 (tpp[IDX(i, j)] - tpp[IDX(i, jp)]) * invdphi);
            vaz[IDX(i, j)] += dVp;
        }
#pragma omp parallel for
    for (int i = 2; i < nr - 1; ++i)
        for (int j = 0; j < ns; ++j) {
            const int jn = j == ns - 1 ? 0 : j + 1;
            const double sigma_avg = 0.5 * (S[IDX(i, j)] + S[IDX(i - 1, j)]);
            const double dVr =
                dt / (sigma_avg)*d->p.radial_viscosity_factor * 2.0 /
                (d->Rmed[i] + d->Rmed[i - 1]) *
                ((d->Rmed[i] * trr[IDX(i, j)] - d->Rmed[i - 1] * trr[IDX(i - 1, j)]) *
                     d->InvDiffRmed[i] +
                 (trp[IDX(i, jn)] - trp[IDX(i, j)]) * invdphi -
                 0.5 * (tpp[IDX(i, j)] + tpp[IDX(i - 1, j)]));
            vr[IDX(i, j)] += dVr;
        }
}

/* thermal_relaxation (beta cooling), SourceEuler.cpp:632-690 */
static void thermal_relaxation(fo_disk *d, double current_time)
{
    const int nr = d->nr, ns = d->ns;
    double *Qm = d->f[FCPT_QMINUS], *E = d->f[FCPT_ENERGY], *S = d->f[FCPT_SIGMA],
           *S0 = d->f[FCPT_SIGMA0], *E0 = d->f[FCPT_ENERGY0];
#pragma omp parallel for
    for (int i = 1; i < nr - 1; ++i)
        for (int j = 0; j < ns; ++j) {
            const double r = d->Rmed[i];
            const double omega_k = omega_kepler(d, r);
            const double e = E[IDX(i, j)];
            const double t_ramp_up = d->p.cooling_beta_ramp_up;
            double beta_inv = 1 / d->p.cooling_beta;
            if (t_ramp_up > 0.0) {
                const double t = current_time;
                const double ramp_factor = 1 - exp(-pow(2 * t / t_ramp_up, 2));
                beta_inv = beta_inv * ramp_factor;
            }
            double delta_E = e;
            if (d->p.cooling_beta_ref == FCPT_BETA_REF_REFERENCE) {
                const double sigma = S[IDX(i, j)];
                const double sigma0 = S0[IDX(i, j)];
                const double e0 = E0[IDX(i, j)];
                delta_E -= e0 / sigma0 * sigma;
            }
            if (d->p.cooling_beta_ref == FCPT_BETA_REF_FLOOR) {
                const double minimum_energy =
                    d->p.t_min * S[IDX(i, j)] / d->p.mu * d->p.R / (d->p.gamma - 1.0);
                delta_E -= minimum_energy;
            }
            const double qminus = delta_E * omega_k * beta_inv;
            Qm[IDX(i, j)] += qminus;
        }
}

/* viscous_heating, SourceEuler.cpp:496-536 */
static void viscous_heating(fo_disk *d)
{
    const int nr = d->nr, ns = d->ns;
    double *Qp = d->f[FCPT_QPLUS], *nu = d->f[FCPT_VISCOSITY], *S = d->f[FCPT_SIGMA],
           *trr = d->f[FCPT_TAU_R_R], *tpp = d->f[FCPT_TAU_PHI_PHI], *trp = d->f[FCPT_TAU_R_PHI],
           *divv = d->f[FCPT_DIV_V];
#pragma omp parallel for
    for (int i = 1; i < nr - 1; ++i)
        for (int j = 0; j < ns; ++j) {
            if (nu[IDX(i, j)] != 0.0) {
                const int jn = j == ns - 1 ? 0 : j + 1;
                const double tau_r_phi = 0.25 * (trp[IDX(i, j)] + trp[IDX(i + 1, j)] +
                                                 trp[IDX(i, jn)] + trp[IDX(i + 1, jn)]);
                double qplus = 1.0 / (2.0 * nu[IDX(i, j)] * S[IDX(i, j)]) *
                               (pow(trr[IDX(i, j)], 2) + 2 * pow(tau_r_phi, 2) +
                                pow(tpp[IDX(i, j)], 2));
                qplus += (2.0 / 9.0) * nu[IDX(i, j)] * S[IDX(i, j)] * pow(divv[IDX(i, j)], 2);
                qplus *= d->p.heating_viscous_factor;
                Qp[IDX(i, j)] += qplus;
            }
        }
}

/* SubStep3, SourceEuler.cpp:859-954 (no surface cooling, no irradiation) */
void fo_substep3(fo_disk *d, double time, double dt)
{
    const int nr = d->nr, ns = d->ns;
    compute_temperature(d);
    memset(d->f[FCPT_QMINUS], 0, sizeof(double) * nr * ns);
    if (d->p.cooling_beta_enabled) thermal_relaxation(d, time);
    memset(d->f[FCPT_QPLUS], 0, sizeof(double) * nr * ns);
    if (d->p.heating_viscous) viscous_heating(d);
    double *Qp = d->f[FCPT_QPLUS], *Qm = d->f[FCPT_QMINUS], *H = d->f[FCPT_SCALE_HEIGHT],
           *S = d->f[FCPT_SIGMA], *E = d->f[FCPT_ENERGY];
    const double sigma_sb = d->p.sigma_sb, c = d->p.c_light, mu = d->p.mu, gamma = d->p.gamma,
                 Rgas = d->p.R;
#pragma omp parallel for
    for (int i = 1; i < nr - 1; ++i)
        for (int j = 0; j < ns; ++j) {
            const double h = H[IDX(i, j)];
            const double sigma = S[IDX(i, j)];
            const double energy = E[IDX(i, j)];
            const double inv_pow4 = pow(mu * (gamma - 1.0) / (Rgas * sigma), 4);
            double alpha = 1.0 + 2.0 * h * 4.0 * sigma_sb / c * inv_pow4 * pow(energy, 3);
            Qp[IDX(i, j)] /= alpha;
            Qm[IDX(i, j)] /= alpha;
            const double Qplus = Qp[IDX(i, j)];
            const double Qminus = Qm[IDX(i, j)];
            double energy_new = energy + dt * (Qplus - Qminus);
            const double SigmaFloor = 10.0 * d->p.sigma0 * d->p.sigma_floor;
            if ((sigma < SigmaFloor)) {
                /* TAU_EFF is never computed without surface cooling: it holds 0 */
                const double tau_eff = 0.0;
                const double e4 = Qplus * tau_eff / (2.0 * sigma_sb);
                const double constant = (Rgas / mu * sigma / (gamma - 1.0));
                const double eq_energy = pow(e4, 1.0 / 4.0) * constant;
                Qm[IDX(i, j)] = Qplus;
                energy_new = eq_energy;
            }
            E[IDX(i, j)] = energy_new;
        }
    assure_temperature_range(d);
    (void)Qm;
}

/* damping_single_inner / _outer, boundary_conditions/damping.cpp:311-427 */
static void damp_one(fo_disk *d, double *X, const double *X0, int vector, double dt)
{
    const int ns = d->ns;
    const double *radius = vector ? d->Rinf : d->Rmed;
    const int size_radial = vector ? d->nr + 1 : d->nr;
    const int max_radial = size_radial - 1;
    const double lin = d->p.damping_inner_limit, lout = d->p.damping_outer_limit;
    const double RMIN = d->rmin, RMAX = d->rmax;
    if ((lin > 1.0) && (radius[0] < RMIN * lin)) {
        const int limit = vector ? d->damp_ids[1] : d->damp_ids[0];
        const double tau = d->p.damping_time_factor * 2.0 * M_PI / omega_kepler(d, RMIN);
#pragma omp parallel for
        for (int i = 0; i <= limit; ++i) {
            const double factor = pow((radius[i] - RMIN * lin) / (RMIN - RMIN * lin), 2);
            const double exp_factor = exp(-dt * factor / tau);
            for (int j = 0; j < ns; ++j) {
                const double x = X[IDX(i, j)];
                const double x0 = X0[IDX(i, j)];
                X[IDX(i, j)] = (x - x0) * exp_factor + x0;
            }
        }
    }
    if ((lout < 1.0) && (radius[max_radial] > RMAX * lout)) {
        const int limit = vector ? d->damp_ids[3] : d->damp_ids[2];
        const double tau = d->p.damping_time_factor * 2.0 * M_PI /
                           omega_kepler(d, d->p.damping_time_radius_outer);
#pragma omp parallel for
        for (int i = limit; i < size_radial; ++i) {
            const double factor = pow((radius[i] - RMAX * lout) / (RMAX - RMAX * lout), 2);
            const double exp_factor = exp(-dt * factor / tau);
            for (int j = 0; j < ns; ++j) {
                const double x = X[IDX(i, j)];
                const double x0 = X0[IDX(i, j)];
                X[IDX(i, j)] = (x - x0) * exp_factor + x0;
            }
        }
    }
}

/* apply_boundary_condition, boundary_conditions.cpp:65-114 with the
   reflecting / zero-gradient / outflow composites and Keplerian v_phi */
void fo_apply_boundary_condition(fo_disk *d, double dt, int final_call)
{
    const int nr = d->nr, ns = d->ns;
    double *S = d->f[FCPT_SIGMA], *E = d->f[FCPT_ENERGY], *vr = d->f[FCPT_VRAD],
           *vaz = d->f[FCPT_VAZI];
    if (final_call && d->p.damping) {
        /* damping(), damping.cpp:754-776: V_RADIAL, V_AZIMUTHAL, SIGMA, ENERGY */
        damp_one(d, vr, d->f[FCPT_VRAD0], 1, dt);
        damp_one(d, vaz, d->f[FCPT_VAZI0], 0, dt);
        damp_one(d, S, d->f[FCPT_SIGMA0], 0, dt);
        damp_one(d, E, d->f[FCPT_ENERGY0], 0, dt);
    }
    /* sigma and energy: zero gradient for every supported composite
       (zero_gradient.cpp:16-24, 48-56) */
    for (int j = 0; j < ns; ++j) {
        S[IDX(0, j)] = S[IDX(1, j)];
        S[IDX(nr - 1, j)] = S[IDX(nr - 2, j)];
    }
    for (int j = 0; j < ns; ++j) {
        E[IDX(0, j)] = E[IDX(1, j)];
        E[IDX(nr - 1, j)] = E[IDX(nr - 2, j)];
    }
    /* v_r (vector grid, max radial index nr) */
    for (int j = 0; j < ns; ++j) {
        if (d->p.bc_inner == FCPT_BC_REFLECTING) { /* reflecting.cpp:15-26 */
            vr[IDX(0, j)] = -vr[IDX(2, j)];
            vr[IDX(1, j)] = 0;
        } else if (d->p.bc_inner == FCPT_BC_ZEROGRADIENT) { /* zero_gradient.cpp:26-33 */
            vr[IDX(0, j)] = vr[IDX(2, j)];
            vr[IDX(1, j)] = vr[IDX(2, j)];
        } else { /* outflow.cpp:14-31 */
            if (vr[IDX(2, j)] > 0.0) {
                vr[IDX(1, j)] = 0.0;
                vr[IDX(0, j)] = 0.0;
            } else {
                vr[IDX(1, j)] = vr[IDX(2, j)];
                vr[IDX(0, j)] = vr[IDX(2, j)];
            }
        }
        if (d->p.bc_outer == FCPT_BC_REFLECTING) { /* reflecting.cpp:28-39 */
            vr[IDX(nr, j)] = -vr[IDX(nr - 2, j)];
            vr[IDX(nr - 1, j)] = 0;
        } else if (d->p.bc_outer == FCPT_BC_ZEROGRADIENT) { /* zero_gradient.cpp:58-65 */
            vr[IDX(nr, j)] = vr[IDX(nr - 2, j)];
            vr[IDX(nr - 1, j)] = vr[IDX(nr - 2, j)];
        } else { /* outflow.cpp:33-57 */
            if (vr[IDX(nr - 2, j)] < 0.0) {
                vr[IDX(nr - 1, j)] = 0.0;
                vr[IDX(nr, j)] = 0.0;
            } else {
                vr[IDX(nr - 1, j)] = vr[IDX(nr - 2, j)];
                vr[IDX(nr, j)] = vr[IDX(nr - 2, j)];
            }
        }
    }
    /* keplerian_azimuthal_inner / _outer, keplerian_azimuthal.cpp:19-61 */
    const double vk_in = sqrt(d->p.G * d->p.hydro_center_mass / d->Rmed[0]);
    const double vk_out = sqrt(d->p.G * d->p.hydro_center_mass / d->Rmed[nr - 1]);
    for (int j = 0; j < ns; ++j) {
        vaz[IDX(0, j)] = 1.0 * vk_in - d->Rmed[0] * d->p.omega_frame;
        vaz[IDX(nr - 1, j)] = 1.0 * vk_out - d->Rmed[nr - 1] * d->p.omega_frame;
    }
}

/* ---------------------------------------------------------------------- */
/* Transport, TransportEuler.cpp                                           */

static double van_leer_lim(double a, double b)
{
    /* TransportEuler.cpp:306-312 */
    if (a * b > 0.0) return 2.0 * a * b / (a + b);
    return 0;
}

/* compute_star_radial, :349-406 */
static void star_radial(fo_disk *d, const double *Q, const double *vr, double *qstar, double dt)
{
    const int nr = d->nr, ns = d->ns;
    double *dq = d->dq;
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            if (i == 0 || i == nr - 1) {
                dq[IDX(i, j)] = 0.0;
            } else {
                const double dqm = (Q[IDX(i, j)] - Q[IDX(i - 1, j)]) * d->InvDiffRmed[i];
                const double dqp = (Q[IDX(i + 1, j)] - Q[IDX(i, j)]) * d->InvDiffRmed[i + 1];
                dq[IDX(i, j)] = van_leer_lim(dqp, dqm);
            }
        }
#pragma omp parallel for
    for (int i = 1; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            if (vr[IDX(i, j)] > 0.0)
                qstar[IDX(i, j)] = Q[IDX(i - 1, j)] + (d->Rmed[i] - d->Rmed[i - 1] -
                                                       vr[IDX(i, j)] * dt) * 0.5 * dq[IDX(i - 1, j)];
            else
                qstar[IDX(i, j)] = Q[IDX(i, j)] - (d->Rmed[i + 1] - d->Rmed[i] +
                                                   vr[IDX(i, j)] * dt) * 0.5 * dq[IDX(i, j)];
        }
    for (int j = 0; j < ns; ++j) qstar[IDX(0, j)] = 0.0;
}

/* VanLeerRadial, :545-619 (mass-flow diagnostics omitted) */
static void van_leer_radial(fo_disk *d, const double *vr, double *Q, double dt)
{
    const int nr = d->nr, ns = d->ns;
    const size_t n = (size_t)nr * ns;
    for (size_t k = 0; k < n; ++k) d->work[k] = Q[k] / d->densint[k];
    star_radial(d, d->work, vr, d->qrstar, dt);
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            const double varq_inf = dt * d->dphi * d->Rinf[i] * d->qrstar[IDX(i, j)] *
                                    d->densstar[IDX(i, j)] * vr[IDX(i, j)];
            const double varq_sup = dt * d->dphi * d->Rsup[i] * d->qrstar[IDX(i + 1, j)] *
                                    d->densstar[IDX(i + 1, j)] * vr[IDX(i + 1, j)];
            Q[IDX(i, j)] += (varq_inf - varq_sup) * d->InvSurf[i];
        }
}

/* ComputeStarTheta, :416-465 */
static void star_theta(fo_disk *d, const double *Q, const double *v, double *qstar, double dt)
{
    const int nr = d->nr, ns = d->ns;
    double *dq = d->dq;
#pragma omp parallel for
    for (int i = 0; i < nr; ++i) {
        const double dxtheta = d->dphi * d->Rmed[i];
        const double invdxtheta = 1.0 / dxtheta;
        for (int j = 0; j < ns; ++j) {
            const int jp = j == ns - 1 ? 0 : j + 1;
            const int jm = j == 0 ? ns - 1 : j - 1;
            const double dqm = (Q[IDX(i, j)] - Q[IDX(i, jm)]);
            const double dqp = (Q[IDX(i, jp)] - Q[IDX(i, j)]);
            dq[IDX(i, j)] = 0.5 * van_leer_lim(dqp, dqm) * invdxtheta;
        }
    }
#pragma omp parallel for
    for (int i = 0; i < nr; ++i) {
        const double dxtheta = d->dphi * d->Rmed[i];
        for (int j = 0; j < ns; ++j) {
            const int jm = j == 0 ? ns - 1 : j - 1;
            const double ksi = v[IDX(i, j)] * dt;
            if (ksi > 0.0)
                qstar[IDX(i, j)] = Q[IDX(i, jm)] + (dxtheta - ksi) * dq[IDX(i, jm)];
            else
                qstar[IDX(i, j)] = Q[IDX(i, j)] - (dxtheta + ksi) * dq[IDX(i, j)];
        }
    }
}

/* VanLeerTheta, :630-664 (FARGO: NoSplitAdvection = NO, so every ring is updated) */
static void van_leer_theta(fo_disk *d, const double *v, double *Q, double dt)
{
    const int nr = d->nr, ns = d->ns;
    const size_t n = (size_t)nr * ns;
    for (size_t k = 0; k < n; ++k) d->work[k] = Q[k] / d->densint[k];
    star_theta(d, d->work, v, d->qrstar, dt);
#pragma omp parallel for
    for (int i = 0; i < nr; ++i) {
        const double dxrad = (d->Rsup[i] - d->Rinf[i]) * dt;
        const double invsurf = d->InvSurf[i];
        for (int j = 0; j < ns; ++j) {
            const int jp = j == ns - 1 ? 0 : j + 1;
            double varq = dxrad * d->qrstar[IDX(i, j)] * d->densstar[IDX(i, j)] * v[IDX(i, j)];
            varq -= dxrad * d->qrstar[IDX(i, jp)] * d->densstar[IDX(i, jp)] * v[IDX(i, jp)];
            Q[IDX(i, j)] += varq * invsurf;
        }
    }
}

static void quantities_advection(fo_disk *d, const double *v, double dt)
{
    /* QuantitiesAdvection, :292-304 */
    const size_t n = (size_t)d->nr * d->ns;
    star_theta(d, d->f[FCPT_SIGMA], v, d->densstar, dt);
    memcpy(d->densint, d->f[FCPT_SIGMA], n * sizeof(double));
    for (int k = 0; k < 4; ++k) van_leer_theta(d, v, d->mom[k], dt);
    if (d->p.adiabatic) van_leer_theta(d, v, d->f[FCPT_ENERGY], dt);
    van_leer_theta(d, v, d->f[FCPT_SIGMA], dt);
}

static void advect_shift(fo_disk *d, double *val)
{
    /* AdvectSHIFT, :238-268 */
    const int nr = d->nr, ns = d->ns;
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            int ji = j - d->nshift[i];
            while (ji < 0) ji += ns;
            while (ji >= ns) ji -= ns;
            d->tmp[IDX(i, j)] = val[IDX(i, ji)];
        }
    memcpy(val, d->tmp, sizeof(double) * nr * ns);
}

void fo_transport(fo_disk *d, double dt)
{
    const int nr = d->nr, ns = d->ns;
    const size_t n = (size_t)nr * ns;
    double *S = d->f[FCPT_SIGMA], *vr = d->f[FCPT_VRAD], *vaz = d->f[FCPT_VAZI],
           *E = d->f[FCPT_ENERGY];
    const double OmegaF = d->p.omega_frame;
    /* compute_momenta_from_velocities, :471-495 */
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            d->mom[0][IDX(i, j)] = S[IDX(i, j)] * vr[IDX(i + 1, j)];
            d->mom[1][IDX(i, j)] = S[IDX(i, j)] * vr[IDX(i, j)];
            const int jn = j == ns - 1 ? 0 : j + 1;
            const double vnext = vaz[IDX(i, jn)];
            const double r = d->Rmed[i];
            d->mom[2][IDX(i, j)] = S[IDX(i, j)] * (vnext + r * OmegaF) * r;
            d->mom[3][IDX(i, j)] = S[IDX(i, j)] * (vaz[IDX(i, j)] + r * OmegaF) * r;
        }
    /* OneWindRad, :138-168 */
    star_radial(d, S, vr, d->densstar, dt);
    memcpy(d->densint, S, n * sizeof(double));
    for (int k = 0; k < 4; ++k) van_leer_radial(d, vr, d->mom[k], dt);
    if (d->p.adiabatic) van_leer_radial(d, vr, E, dt);
    van_leer_radial(d, vr, S, dt);
    /* OneWindTheta, :270-289 */
#pragma omp parallel for
    for (int i = 0; i < nr; ++i) { /* compute_average_azimuthal_velocity, :174-190 */
        double sum = 0.0;
        for (int j = 0; j < ns; ++j) sum += vaz[IDX(i, j)];
        d->vmean[i] = sum / (double)ns;
    }
#pragma omp parallel for
    for (int i = 0; i < nr; ++i) /* compute_residual_velocity, :195-205 */
        for (int j = 0; j < ns; ++j) d->vres[IDX(i, j)] = vaz[IDX(i, j)] - d->vmean[i];
    {
        /* ComputeConstantResidual, :207-236 (FARGO) */
        const double invdt = 1.0 / dt;
#pragma omp parallel for
        for (int i = 0; i < nr; i++) {
            const double Ntilde = d->vmean[i] * d->InvRmed[i] * dt * d->invdphi;
            const double Nround = floor(Ntilde + 0.5);
            d->nshift[i] = (int)Nround;
            for (int j = 0; j < ns; j++)
                vaz[IDX(i, j)] = (Ntilde - Nround) * d->Rmed[i] * invdt * d->dphi;
        }
    }
    quantities_advection(d, d->vres, dt);
    quantities_advection(d, vaz, dt);
    for (int k = 0; k < 4; ++k) advect_shift(d, d->mom[k]);
    if (d->p.adiabatic) advect_shift(d, E);
    advect_shift(d, S);
    /* compute_velocities_from_momenta, :498-535 */
#pragma omp parallel for
    for (int i = 0; i < nr; ++i)
        for (int j = 0; j < ns; ++j) {
            const int jm = j == 0 ? ns - 1 : j - 1;
            if (i == 0) {
                vr[IDX(i, j)] = 0.0;
            } else {
                vr[IDX(i, j)] = (d->mom[0][IDX(i - 1, j)] + d->mom[1][IDX(i, j)]) /
                                (S[IDX(i - 1, j)] + S[IDX(i, j)]);
            }
            vaz[IDX(i, j)] = (d->mom[2][IDX(i, jm)] + d->mom[3][IDX(i, j)]) /
                                 (S[IDX(i, jm)] + S[IDX(i, j)]) * d->InvRmed[i] -
                             d->Rmed[i] * OmegaF;
        }
    /* assure_minimum_value(Sigma), SourceEuler.cpp:102-131 */
    const double floor_value = d->p.sigma_floor * d->p.sigma0;
    for (size_t k = 0; k < n; ++k)
        if (S[k] < floor_value) S[k] = floor_value;
    if (d->p.adiabatic) assure_temperature_range(d);
}

/* recalculate_derived_disk_quantities, SourceEuler.cpp:225-249 */
void fo_recalculate_derived(fo_disk *d)
{
    if (!d->p.adiabatic) {
        compute_pressure(d);
    } else {
        compute_temperature(d);
        compute_sound_speed(d);
        compute_scale_height(d);
        compute_pressure(d);
    }
    update_viscosity(d);
}

void fo_compute_pressure(fo_disk *d) { compute_pressure(d); }

/* the source part of one LeapFrog kick (simulation.cpp:319-340, 364-385) */
static void leapfrog_kick(fo_disk *d, double kick_time, double frog_dt)
{
    fo_update_with_sourceterms(d, frog_dt);
    fo_artificial_viscosity(d, frog_dt);
    fo_recalculate_viscosity(d);
    fo_viscous_stress_tensor(d);
    fo_update_velocities_with_viscosity(d, frog_dt);
    if (d->p.adiabatic) fo_substep3(d, kick_time, frog_dt);
}

/* step_LeapFrog gas part, simulation.cpp:276-459: kick dt/2 with the bodies
   of the step parameters, drift (boundaries, transport over dt), kick dt/2
   with the midstep bodies after compute_pressure, final boundaries, derived */
void fo_step_leapfrog(fo_disk *d, double step_dt)
{
    const double frog_dt = step_dt / 2;
    const double start_time = d->time;
    const double midstep_time = d->time + frog_dt;
    const double end_time = d->time + step_dt;
    fo_nbody_potential(d);
    leapfrog_kick(d, start_time, frog_dt);
    fo_apply_boundary_condition(d, 0.0, 0);
    fo_transport(d, step_dt);
    const fcpt_step_params first = d->sp;
    if (d->has_mid) d->sp = d->sp_mid;
    fo_nbody_potential(d);
    d->sp = first;
    compute_pressure(d);
    leapfrog_kick(d, midstep_time, frog_dt);
    d->time = end_time;
    fo_apply_boundary_condition(d, step_dt, 1);
    fo_recalculate_derived(d);
}

/* step_Euler gas part, simulation.cpp:148-267 */
void fo_step_euler(fo_disk *d, double dt)
{
    fo_nbody_potential(d);
    fo_update_with_sourceterms(d, dt);
    fo_artificial_viscosity(d, dt);
    fo_recalculate_viscosity(d);
    fo_viscous_stress_tensor(d);
    fo_update_velocities_with_viscosity(d, dt);
    if (d->p.adiabatic) fo_substep3(d, d->time, dt);
    fo_apply_boundary_condition(d, dt, 0);
    fo_transport(d, dt);
    d->time += dt;
    fo_apply_boundary_condition(d, dt, 1);
    fo_recalculate_derived(d);
}

/* one iteration of sim::run, simulation.cpp:515-556 */
double fo_run_step(fo_disk *d)
{
    const double cfl_dt = fo_calculate_timestep(d);
    const double time_next_monitor = (d->n_monitor + 1) * d->p.monitor_timestep;
    const double time_left_till_write = time_next_monitor - d->time;
    const int overshoot = cfl_dt > time_left_till_write;
    const double dt_stretch_factor = 0.05;
    const int almost_there = time_left_till_write < cfl_dt * (1 + dt_stretch_factor);
    const double step_dt = (overshoot || almost_there) ? time_left_till_write : cfl_dt;
    if (d->p.integrator == FCPT_INTEGRATOR_LEAPFROG) fo_step_leapfrog(d, step_dt);
    else fo_step_euler(d, step_dt);
    if (fabs(time_next_monitor - d->time) < 1e-6 * cfl_dt) d->n_monitor++;
    return step_dt;
}

/* gas_total_mass / gas_angular_momentum (quantities.cpp:51-78, 242-279), summed
   ring by ring with Neumaier compensation */
void fo_totals(fo_disk *d, double radius, double *mass, double *angmom)
{
    const int nr = d->nr, ns = d->ns;
    const double *S = d->f[FCPT_SIGMA], *vaz = d->f[FCPT_VAZI];
    double m = 0.0, mc = 0.0, l = 0.0, lc = 0.0;
    for (int i = 1; i < nr - 1; ++i) {
        if (!(d->Rmed[i] <= radius)) continue;
        for (int j = 0; j < ns; ++j) {
            const int jm = j == 0 ? ns - 1 : j - 1;
            const double dm = d->Surf[i] * S[IDX(i, j)];
            const double dl = d->Surf[i] * 0.5 * (S[IDX(i, j)] + S[IDX(i, jm)]) * d->Rmed[i] *
                              (vaz[IDX(i, j)] + d->p.omega_frame * d->Rmed[i]);
            double t = m + dm;
            mc += fabs(m) >= fabs(dm) ? (m - t) + dm : (dm - t) + m;
            m = t;
            t = l + dl;
            lc += fabs(l) >= fabs(dl) ? (l - t) + dl : (dl - t) + l;
            l = t;
        }
    }
    *mass = m + mc;
    *angmom = l + lc;
}

/* ComputeAverageDensity (Pframeforce.cpp:174-187) into a ring array, then
   ComputeDiskOnPlanetAccel (Force.cpp:23-122) for every body, with the
   reference's sequential order of a one-thread run: for each ring of
   radial_first_active .. radial_active_size-1, each sector in order, the
   inner (Rmed < a) or outer accumulator; compute_smoothing, Force.cpp:145-159 */
void fo_disk_on_planet_accel(fo_disk *d, const fcpt_force_body *bodies, int n_bodies, int flags,
                             double *accel_xy)
{
    const int nr = d->nr, ns = d->ns;
    const double *S = d->f[FCPT_SIGMA], *H = d->f[FCPT_SCALE_HEIGHT];
    const double G = d->p.G;
    double *sigma1d = calloc((size_t)nr, sizeof(double));
    if (flags & FCPT_FORCE_CORRECT_SELFGRAVITY) {
        for (int i = 1; i < nr - 1; ++i) {
            double sum = 0;
            for (int j = 0; j < ns; ++j) sum += S[IDX(i, j)];
            sigma1d[i] = sum / ns;
        }
    }
    for (int nb = 0; nb < n_bodies; ++nb) {
        const double x = bodies[nb].x, y = bodies[nb].y, a = bodies[nb].r;
        const double r_sm = bodies[nb].r_cubic;
        double axi = 0.0, ayi = 0.0, axo = 0.0, ayo = 0.0;
        for (int i = 1; i < nr - 1; ++i) {
            for (int j = 0; j < ns; ++j) {
                double smooth;
                if ((flags & FCPT_FORCE_NO_STAR_SMOOTHING) && nb == 0) {
                    smooth = 0;
                } else if (flags & FCPT_FORCE_SMOOTHING_PLANETLOC) {
                    const double scale_height =
                        d->p.aspect_ratio * pow(a, 1 + d->p.flaring_index);
                    smooth = d->p.thickness_smoothing * scale_height;
                } else {
                    smooth = d->p.thickness_smoothing * H[IDX(i, j)];
                }
                const double xc = d->Rmed[i] * d->cosphi[j];
                const double yc = d->Rmed[i] * d->sinphi[j];
                double cell_sigma = S[IDX(i, j)];
                if (flags & FCPT_FORCE_CORRECT_SELFGRAVITY) cell_sigma -= sigma1d[i];
                const double cellmass = d->Surf[i] * cell_sigma;
                const double dx = xc - x;
                const double dy = yc - y;
                const double dist_2 = pow(dx, 2) + pow(dy, 2);
                const double dist_sm_2 = dist_2 + pow(smooth, 2);
                const double dist_sm = sqrt(dist_sm_2);
                const double dist_sm_3 = dist_sm_2 * dist_sm;
                const double inv_dist_sm_3 = 1.0 / dist_sm_3;
                double smooth_factor_klahr = 1.0;
                if (r_sm > 0.0 && dist_sm < r_sm) {
                    smooth_factor_klahr =
                        -(3.0 * pow(dist_sm / r_sm, 4.0) - 4.0 * pow(dist_sm / r_sm, 3.0));
                }
                if (d->Rmed[i] < a) {
                    axi += G * cellmass * dx * inv_dist_sm_3 * smooth_factor_klahr;
                    ayi += G * cellmass * dy * inv_dist_sm_3 * smooth_factor_klahr;
                } else {
                    axo += G * cellmass * dx * inv_dist_sm_3 * smooth_factor_klahr;
                    ayo += G * cellmass * dy * inv_dist_sm_3 * smooth_factor_klahr;
                }
            }
        }
        accel_xy[2 * nb] = axi + axo;
        accel_xy[2 * nb + 1] = ayi + ayo;
    }
    free(sigma1d);
}
