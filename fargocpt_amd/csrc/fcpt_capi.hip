// C ABI of libfcpt_hip.so (include/fcpt.h): context management, state
// transfer and one entry point per reference hydro-step function.  All device
// work of a context is enqueued on the context's own HIP stream (default
// priority); functions that hand a value back to the host synchronise it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "fcpt_device.hpp"
#include "fcpt_kernels.hpp"

using namespace fcpt;

// split storage: steps of fcpt_run_steps between two moves of the ring
// backgrounds unless the caller sets another interval
constexpr int kDefaultRebaseInterval = 1000;

struct fcpt_ctx {
    fcpt_disk_desc desc{};
    KParams P{};
    hipStream_t stream = nullptr;
    int device = 0;
    std::vector<void *> allocs;
    bool fused = true;  // fused step kernels (FCPT_PARAM_FUSED_STEP)
    DevScalars *s_dev = nullptr;
    DevScalars *s_host = nullptr;  // pinned mirror for read-back
    double *totals_ring = nullptr;
    double *totals_out = nullptr;
    double *totals_host = nullptr;
    double *force_sig1d = nullptr;  // ring means of Sigma for the disk-force reduction
    double *force_part = nullptr;   // per-block partial sums
    double *force_out = nullptr;
    double *force_host = nullptr;
    hipGraphExec_t step_graph = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    // host copies of the per-ring metric (split storage derives its backgrounds from them)
    std::vector<double> Rinf, Rmed, Surf, InvRmed, InvDiffRmed, InvDiffRsup, InvDiffRsupRb;
    // split storage: host copies of the backgrounds last set from the host, and
    // a device ring array for re-expressing a field about a new background
    std::vector<double> hbg[FCPT_NUM_FIELDS];
    double *bg_tmp = nullptr;
    // fused FP64 adiabatic steps leave c_s, nu and T to be formed (lazy_derived)
    bool derived_stale = false;
    // split storage: steps between background moves (FCPT_PARAM_REBASE_INTERVAL;
    // 0 = only on request) and steps since the last one
    int rebase_interval = 0;
    int steps_since_rebase = 0;
};

namespace {

int fail(fcpt_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIP_TRY(ctx, expr)                                                            \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail((ctx), FCPT_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH_CHECK(ctx)                                                             \
    do {                                                                              \
        hipError_t e_ = hipGetLastError();                                            \
        if (e_ != hipSuccess)                                                         \
            return fail((ctx), FCPT_ERR_DEVICE, "kernel launch: %s", hipGetErrorString(e_)); \
    } while (0)

#define CTX_CHECK(ctx)                                   \
    do {                                                 \
        if (!(ctx)) return FCPT_ERR_ARG;                 \
        hipError_t e_ = hipSetDevice((ctx)->device);     \
        if (e_ != hipSuccess)                            \
            return fail((ctx), FCPT_ERR_DEVICE, "hipSetDevice: %s", hipGetErrorString(e_)); \
    } while (0)

// When the fused Euler step of an adiabatic disk runs the kick and viscous
// walks, the step forms c_s, nu and T where it needs them
// (KParams::lazy_derived); in split storage the kick walk also forms the
// pressure from E and passes E to the viscous walk through its partner buffer
// (KParams::e_partner), so the end of the step does not store the pressure
// (which saves a decode and an encode per cell where the split kernels are
// bound by their arithmetic; in FP64 storage, bound by the bytes, measured
// 0.5 % slower, so not used there).  Afterwards those fields are stale until
// sync_derived stores them (the same expressions on the stored E and Sigma).
static bool lazy_applies(const fcpt_ctx *ctx)
{
    const KParams &P = ctx->P;
    return ctx->fused && P.adiabatic && P.integrator == FCPT_INTEGRATOR_EULER && kick_walk_applies(P);
}
static KParams step_view(const fcpt_ctx *ctx)
{
    KParams P = ctx->P;
    P.e_partner = lazy_applies(ctx) && P.split ? 1 : 0;
    P.lazy_derived = lazy_applies(ctx) ? 1 : 0;
    return P;
}
static int sync_derived(fcpt_ctx *ctx)
{
    if (!ctx->derived_stale) return FCPT_OK;
    launch_derived(ctx->P, ctx->stream);
    ctx->derived_stale = false;
    return FCPT_OK;
}
// entry points that read or expose the fields: the context, with the derived
// fields stored
#define CTX_SYNCED(ctx)                        \
    do {                                       \
        CTX_CHECK(ctx);                        \
        const int rc_sync_ = sync_derived(ctx); \
        if (rc_sync_) return rc_sync_;         \
    } while (0)

int dev_alloc_bytes(fcpt_ctx *c, void **p, size_t bytes)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes + 256);
    if (e != hipSuccess)
        return fail(c, FCPT_ERR_DEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    c->allocs.push_back(q);
    *p = q;
    return FCPT_OK;
}

template <typename T>
int dev_alloc(fcpt_ctx *c, T **p, size_t count)
{
    void *q = nullptr;
    const int rc = dev_alloc_bytes(c, &q, count * sizeof(T));
    *p = static_cast<T *>(q);
    return rc;
}

size_t field_rows(int field, int nr)
{
    return (field == FCPT_VRAD || field == FCPT_VRAD0 || field == FCPT_TAU_R_PHI) ? (size_t)nr + 1
                                                                                : (size_t)nr;
}

// log-mesh ring lookups with the reference's NDEBUG clamps
// (find_cell_id.cpp:218-224, 259-264, 107-215; damping.cpp:321-327, 377-383)
int rmed_id_log(const fcpt_disk_desc &d, double r)
{
    const double g = std::pow(d.rmax / d.rmin, 1.0 / ((double)d.nr - 2.0));
    const double opt = 3.0 / 2.0 / d.rmin * (1 - std::pow(g, 2.0)) / (1 - std::pow(g, 3.0));
    const double inv_log = 1.0 / std::log(g);
    int id = (int)std::floor(std::log(r * opt) * inv_log) + 1;
    if (id < 0) id = 0;
    if (id > d.nr - 1) id = d.nr - 1;
    return id;
}
int rinf_id_log(const fcpt_disk_desc &d, double r)
{
    const double g = std::pow(d.rmax / d.rmin, 1.0 / ((double)d.nr - 2.0));
    const double inv_log = 1.0 / std::log(g);
    int id = (int)std::floor(std::log(r / d.rmin) * inv_log) + 1;
    if (id < 0) id = 0;
    if (id > d.nr - 1) id = d.nr - 1;
    return id;
}
int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// The radial profiles of the reference's profile initial condition, one
// function of R each, in the reference's operand order so that they round
// the same way (host libm, no contraction): Sigma (init.cpp:946-957 and the
// cutoffs :1047-1148; get_sigma, viscous_radial_speed.cpp:95-118), the
// pressure-supported smoothed v_phi (Theo.cpp:131-181), the viscosity of the
// viscous-speed estimate (get_nu2, viscous_radial_speed.cpp:38-93) and the
// numerical viscous radial speed (:120-195).
struct DiskProfile {
    const fcpt_disk_desc &d;
    const fcpt_ic_desc &ic;

    static double cut_out(double point, double width, double x) { return 1.0 / (1.0 + std::exp((x - point) / width)); }
    static double cut_in(double point, double width, double x) { return 1.0 / (1.0 + std::exp((point - x) / width)); }
    static double max_(double a, double b) { return (a < b) ? b : a; }
    static double min_(double a, double b) { return (b < a) ? b : a; }
    double floor_value() const { return d.sigma_floor * d.sigma0; }

    double cutoff(double R) const
    {
        double c = 1.0;
        if (ic.cutoff_outer) c *= cut_out(ic.cutoff_outer_point, ic.cutoff_outer_width, R);
        if (ic.cutoff_inner) c *= cut_in(ic.cutoff_inner_point, ic.cutoff_inner_width, R);
        return c;
    }
    // the grid's Sigma: power law, floor, then each cutoff with the floor again
    double sigma_cell(double R) const
    {
        double s = max_(d.sigma0 * std::pow(R, -ic.sigma_slope), floor_value());
        if (ic.cutoff_outer) s = max_(s * cut_out(ic.cutoff_outer_point, ic.cutoff_outer_width, R), floor_value());
        if (ic.cutoff_inner) s = max_(s * cut_in(ic.cutoff_inner_point, ic.cutoff_inner_width, R), floor_value());
        return s;
    }
    // the viscous-speed estimate's Sigma: both cutoffs, then the floor once
    double sigma_model(double R) const
    {
        double s = d.sigma0 * std::pow(R, -ic.sigma_slope);
        if (ic.cutoff_outer) s *= cut_out(ic.cutoff_outer_point, ic.cutoff_outer_width, R);
        if (ic.cutoff_inner) s *= cut_in(ic.cutoff_inner_point, ic.cutoff_inner_width, R);
        return max_(s, floor_value());
    }
    double energy(double R) const
    {
        return 1.0 / (d.gamma - 1.0) * d.sigma0 * std::pow(d.aspect_ratio, 2) *
               std::pow(R, -ic.sigma_slope - 1.0 + 2.0 * d.flaring_index) * d.G * d.hydro_center_mass;
    }
    double energy_floor(double sig) const { return d.t_min * sig / d.mu * d.R / (d.gamma - 1.0); }
    double vphi(double R) const
    {
        const double F = d.flaring_index;
        const double h = d.aspect_ratio * std::pow(R, F);
        const double he = h * d.thickness_smoothing;
        const double smoothing = (1.0 + (F + 1.0) * std::pow(he, 2)) / std::pow(std::sqrt(1 + std::pow(he, 2)), 3);
        const double pressure = (2.0 * F - 1.0 - ic.sigma_slope) * std::pow(h, 2);
        const double vk_2 = d.G * d.hydro_center_mass / R;
        return std::sqrt(vk_2 * (smoothing + pressure));
    }
    double nu(double R, double sig) const
    {
        const double v_k = std::sqrt(d.G * d.hydro_center_mass / R);
        const double h = d.aspect_ratio * std::pow(R, d.flaring_index);
        if (!d.adiabatic) return d.alpha * (h * v_k) * (h * R);
        const double g = d.gamma;
        double e = cutoff(R) * 1.0 / (g - 1.0) * sig * std::pow(h * v_k, 2);
        e = max_(e, d.t_min * sig / d.mu * d.R / (g - 1.0));
        e = min_(e, d.t_max * sig / d.mu * d.R / (g - 1.0));
        const double cs_adb = std::sqrt(g * (g - 1.0) * e / sig);
        const double cs_iso = std::sqrt((g - 1.0) * e / sig);
        return d.alpha * cs_adb * (cs_iso / (v_k / R));
    }
    double omega(double r) const { return vphi(r) / r; }
    double r2_omega(double r) const { return std::pow(r, 2) * omega(r); }
    // five-point derivative with the step 8e-4 r
    template <class Fn>
    static double derivative(double r, Fn f)
    {
        const double h = 8.0e-4 * r;
        const double t = -1.0 * f(r + 2.0 * h) + 8.0 * f(r + h) + -8.0 * f(r - h) + 1.0 * f(r - 2.0 * h);
        return t / (12.0 * h);
    }
    double torque_flux(double r) const  // nu Sigma r^3 d(omega)/dr
    {
        const double dw = derivative(r, [this](double x) { return omega(x); });
        const double sig = sigma_model(r);
        return nu(r, sig) * sig * std::pow(r, 3) * dw;
    }
    double vr_viscous(double r) const
    {
        const double num = 1.0 / r * derivative(r, [this](double x) { return torque_flux(x); });
        const double den = sigma_model(r) * derivative(r, [this](double x) { return r2_omega(x); });
        return num / den;
    }
};

}  // namespace

extern "C" {

const char *fcpt_last_error(const fcpt_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fcpt_create(const fcpt_disk_desc *desc, fcpt_ctx **out)
{
    if (!desc || !out) return FCPT_ERR_ARG;
    *out = nullptr;
    if (desc->nr < 8 || desc->nr > 16384 || desc->nphi < 4 || desc->nphi > 262144 || !desc->radii)
        return FCPT_ERR_ARG;
    if (desc->precision != FCPT_PREC_FP64 && desc->precision != FCPT_PREC_SPLIT32) return FCPT_ERR_ARG;
    if (desc->integrator != FCPT_INTEGRATOR_EULER && desc->integrator != FCPT_INTEGRATOR_LEAPFROG)
        return FCPT_ERR_ARG;
    if (desc->potential_smoothing & ~(FCPT_SMOOTHING_NO_STAR | FCPT_SMOOTHING_PLANETLOC)) return FCPT_ERR_ARG;
    for (int i = 0; i + 1 < desc->nr + 2; ++i)
        if (!(desc->radii[i + 1] > desc->radii[i]) || !(desc->radii[i] > 0.0)) return FCPT_ERR_ARG;
    fcpt_ctx *c = new fcpt_ctx();
    c->desc = *desc;
    c->desc.radii = nullptr;
    hipError_t e = hipGetDevice(&c->device);
    if (e != hipSuccess) {
        delete c;
        return FCPT_ERR_DEVICE;
    }
    const int nr = desc->nr, ns = desc->nphi;
    KParams &P = c->P;
    P.nr = nr;
    P.ns = ns;
    P.split = desc->precision == FCPT_PREC_SPLIT32 ? 1 : 0;
    P.split_rounding = P.split;  // stochastic rounding of the split deviations by default
    c->rebase_interval = P.split ? kDefaultRebaseInterval : 0;
    P.radial_tile = 0;
    P.radial_waves = 3;
    P.visc_tile = 0;
    P.kick_tile = 0;
    P.lazy_derived = 0;
    P.e_partner = 0;
    P.ealt = nullptr;
    init_launch_fields(P);
    // Interpret.cpp:230-231
    P.dphi = 2.0 * M_PI / (double)ns;
    P.invdphi = (double)ns / (2.0 * M_PI);
    P.rmin = desc->rmin;
    P.rmax = desc->rmax;
    P.adiabatic = desc->adiabatic;
    P.integrator = desc->integrator;
    P.artvisc = desc->artificial_viscosity;
    P.artvisc_dissipation = desc->artvisc_dissipation;
    P.heating_viscous = desc->heating_viscous;
    P.cooling_beta_enabled = desc->cooling_beta_enabled;
    P.cooling_beta_ref = desc->cooling_beta_ref;
    P.bc_inner = desc->bc_inner;
    P.bc_outer = desc->bc_outer;
    P.damping = desc->damping;
    P.G = desc->G;
    P.Rgas = desc->R;
    P.sigma_sb = desc->sigma_sb;
    P.c_light = desc->c_light;
    P.mstar = desc->hydro_center_mass;
    P.mu = desc->mu;
    P.gamma = desc->gamma;
    P.aspect_ratio = desc->aspect_ratio;
    P.flaring = desc->flaring_index;
    P.alpha = desc->alpha;
    P.thickness_smoothing = desc->thickness_smoothing;
    P.cfl = desc->cfl;
    P.cfl_max_var = desc->cfl_max_var;
    P.heating_cooling_cfl_limit = desc->heating_cooling_cfl_limit;
    P.sigma0 = desc->sigma0;
    P.sigma_floor = desc->sigma_floor;
    P.t_min = desc->t_min;
    P.t_max = desc->t_max;
    P.artvisc_factor = desc->artvisc_factor;
    P.heating_viscous_factor = desc->heating_viscous_factor;
    P.cooling_beta = desc->cooling_beta;
    P.cooling_beta_ramp_up = desc->cooling_beta_ramp_up;
    P.radial_viscosity_factor = desc->radial_viscosity_factor;
    P.damping_inner_limit = desc->damping_inner_limit;
    P.damping_outer_limit = desc->damping_outer_limit;
    P.damping_time_factor = desc->damping_time_factor;
    P.damping_time_radius_outer = desc->damping_time_radius_outer;
    P.omega_frame = desc->omega_frame;
    P.monitor_timestep = desc->monitor_timestep;
    int ids[4] = {desc->damp_inner_scalar, desc->damp_inner_vector, desc->damp_outer_scalar,
                  desc->damp_outer_vector};
    if (ids[0] < 0) ids[0] = clampi(rmed_id_log(*desc, desc->rmin * desc->damping_inner_limit), nr - 1);
    if (ids[1] < 0) ids[1] = clampi(rinf_id_log(*desc, desc->rmin * desc->damping_inner_limit), nr);
    if (ids[2] < 0)
        ids[2] = clampi(rmed_id_log(*desc, desc->rmax * desc->damping_outer_limit) + 1, nr - 1);
    if (ids[3] < 0)
        ids[3] = clampi(rinf_id_log(*desc, desc->rmax * desc->damping_outer_limit) + 1, nr);
    memcpy(P.damp_ids, ids, sizeof(ids));
    if (desc->damping && (ids[0] >= ids[2] || ids[1] >= ids[3])) {
        delete c;
        return FCPT_ERR_ARG;  // inner and outer damping zones overlap
    }

    // per-ring metric arrays, init_radialarrays (init.cpp:188-225), in FP64 on the host
    const int n1 = nr + 2;
    std::vector<double> Rinf(n1, 0), Rsup(n1, 0), Rmed(n1, 0), Surf(n1, 0), InvSurf(n1, 0),
        InvRmed(n1, 0), InvRinf(n1, 0), InvDiffRmed(n1, 0), InvDiffRsup(n1, 0),
        InvDiffRsupRb(n1, 0), cosphi(ns), sinphi(ns);
    for (int i = 0; i < nr + 1; ++i) {
        Rinf[i] = desc->radii[i];
        Rsup[i] = desc->radii[i + 1];
        Rmed[i] = 2.0 / 3.0 * (std::pow(Rsup[i], 3) - std::pow(Rinf[i], 3));
        Rmed[i] = Rmed[i] / (std::pow(Rsup[i], 2) - std::pow(Rinf[i], 2));
        Surf[i] = M_PI * (std::pow(Rsup[i], 2) - std::pow(Rinf[i], 2)) / (double)ns;
        InvRmed[i] = 1.0 / Rmed[i];
        InvSurf[i] = 1.0 / Surf[i];
        InvDiffRsup[i] = 1.0 / (Rsup[i] - Rinf[i]);
        InvDiffRsupRb[i] = 1.0 / ((Rsup[i] - Rinf[i]) * Rmed[i]);
        InvRinf[i] = 1.0 / Rinf[i];
    }
    for (int i = 1; i < nr + 1; ++i) InvDiffRmed[i] = 1.0 / (Rmed[i] - Rmed[i - 1]);
    c->Rinf = Rinf;
    c->Rmed = Rmed;
    c->Surf = Surf;
    c->InvRmed = InvRmed;
    c->InvDiffRmed = InvDiffRmed;
    c->InvDiffRsup = InvDiffRsup;
    c->InvDiffRsupRb = InvDiffRsupRb;
    for (int j = 0; j < ns; ++j) {
        // cell centres, SideEuler.cpp:57-66
        cosphi[j] = std::cos(P.dphi * (double)j);
        sinphi[j] = std::sin(P.dphi * (double)j);
    }

    int rc = FCPT_OK;
    auto upload = [&](const double **dst, const std::vector<double> &src) -> int {
        double *d = nullptr;
        int r = dev_alloc(c, &d, src.size());
        if (r) return r;
        hipError_t e2 = hipMemcpy(d, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e2 != hipSuccess) return fail(c, FCPT_ERR_DEVICE, "hipMemcpy: %s", hipGetErrorString(e2));
        *dst = d;
        return FCPT_OK;
    };
    if ((rc = upload(&P.Rinf, Rinf)) || (rc = upload(&P.Rsup, Rsup)) || (rc = upload(&P.Rmed, Rmed)) ||
        (rc = upload(&P.Surf, Surf)) || (rc = upload(&P.InvSurf, InvSurf)) ||
        (rc = upload(&P.InvRmed, InvRmed)) || (rc = upload(&P.InvRinf, InvRinf)) ||
        (rc = upload(&P.InvDiffRmed, InvDiffRmed)) || (rc = upload(&P.InvDiffRsup, InvDiffRsup)) ||
        (rc = upload(&P.InvDiffRsupRb, InvDiffRsupRb)) || (rc = upload(&P.cosphi, cosphi)) ||
        (rc = upload(&P.sinphi, sinphi))) {
        fcpt_destroy(c);
        return rc;
    }
    const size_t ncell = (size_t)nr * ns;
    const size_t esz = element_size(P);
    for (int k = 0; k < FCPT_NUM_FIELDS && !rc; ++k)
        rc = dev_alloc_bytes(c, &P.f[k], field_rows(k, nr) * ns * field_element_size(P, k));
    if (!rc) rc = dev_alloc_bytes(c, &P.valt[0], field_rows(FCPT_VRAD, nr) * ns * esz);
    if (!rc) rc = dev_alloc_bytes(c, &P.valt[1], field_rows(FCPT_VAZI, nr) * ns * esz);
    if (!rc && P.adiabatic && P.split) rc = dev_alloc_bytes(c, &P.ealt, field_rows(FCPT_ENERGY, nr) * ns * esz);
    for (int k = 0; k < 4 && !rc; ++k) rc = dev_alloc_bytes(c, &P.mom[k], ncell * sizeof(double));  // FP64 in both modes
    for (int k = 0; k < 4 && !rc; ++k) rc = dev_alloc_bytes(c, &P.mom_b[k], ncell * esz);
    // the density and energy between the transport sweeps stay FP64 in both
    // storage modes: they are rounded once per step, on the final store
    if (!rc) rc = dev_alloc_bytes(c, &P.sig_b, ncell * sizeof(double));
    if (!rc) rc = dev_alloc_bytes(c, &P.en_b, ncell * sizeof(double));
    if (!rc) rc = dev_alloc_bytes(c, &P.sig_c, ncell * sizeof(double));
    if (!rc) rc = dev_alloc_bytes(c, &P.en_c, ncell * sizeof(double));
    if (P.split) {
        for (int k = 0; k < FCPT_NUM_FIELDS && !rc; ++k) rc = dev_alloc(c, &P.bg[k], (size_t)nr + 2);
        for (int k = 0; k < 4 && !rc; ++k) rc = dev_alloc(c, &P.bgm[k], (size_t)nr + 2);
        if (!rc) rc = dev_alloc(c, &c->bg_tmp, (size_t)nr + 2);
        for (int k = 0; k < FCPT_NUM_FIELDS; ++k) c->hbg[k].assign((size_t)nr + 2, 0.0);
    }
    if (!rc) rc = dev_alloc(c, &P.iso_cs, (size_t)nr + 2);
    if (!rc) rc = dev_alloc(c, &P.iso_h, (size_t)nr + 2);
    if (!rc) rc = dev_alloc(c, &P.iso_nu, (size_t)nr + 2);
    if (!rc) rc = dev_alloc(c, &P.ring_tab, (size_t)kRingTabs * ((size_t)nr + 2));
    if (!rc) rc = dev_alloc(c, &P.vmean, (size_t)nr + 1);
    if (!rc) rc = dev_alloc(c, &P.vuni, (size_t)nr + 1);
    if (!rc) rc = dev_alloc(c, &P.nshift, (size_t)nr + 1);
    // CFL slots: per block of k_cfl_cells (batches) or per ring of k_cfl_rings
    if (!rc) rc = dev_alloc(c, &P.ring_min, (size_t)(nr > (int)kCflMaxBlocks ? nr : (int)kCflMaxBlocks));
    if (!rc) rc = dev_alloc(c, &c->s_dev, 1);
    if (!rc) rc = dev_alloc(c, &c->totals_ring, 4 * (size_t)nr);
    if (!rc) rc = dev_alloc(c, &c->totals_out, 2);
    if (!rc) rc = dev_alloc(c, &c->force_sig1d, (size_t)nr + 2);
    if (!rc) rc = dev_alloc(c, &c->force_part, (size_t)kCflMaxBlocks * 4 * kForceMaxBodies);
    if (!rc) rc = dev_alloc(c, &c->force_out, (size_t)4 * kForceMaxBodies);
    if (rc) {
        fcpt_destroy(c);
        return rc;
    }
    P.s = c->s_dev;
    e = hipHostMalloc((void **)&c->s_host, sizeof(DevScalars), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void **)&c->totals_host, 2 * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess)
        e = hipHostMalloc((void **)&c->force_host, 4 * kForceMaxBodies * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        fcpt_destroy(c);
        return FCPT_ERR_DEVICE;
    }
    // zero every field (the reference's t_polargrid::set_size clears its arrays);
    // split storage: zero backgrounds and zero deviations
    for (int k = 0; k < FCPT_NUM_FIELDS; ++k)
        e = e == hipSuccess ? hipMemset(P.f[k], 0, field_rows(k, nr) * ns * field_element_size(P, k)) : e;
    for (int k = 0; P.split && k < FCPT_NUM_FIELDS; ++k)
        e = e == hipSuccess ? hipMemset(P.bg[k], 0, ((size_t)nr + 2) * sizeof(double)) : e;
    for (int k = 0; P.split && k < 4; ++k)
        e = e == hipSuccess ? hipMemset(P.bgm[k], 0, ((size_t)nr + 2) * sizeof(double)) : e;
    std::memset(c->s_host, 0, sizeof(DevScalars));
    if (e == hipSuccess) e = hipMemcpy(c->s_dev, c->s_host, sizeof(DevScalars), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        launch_ring_tables(P, c->stream);
        if (!P.adiabatic) launch_iso_rings(P, c->stream);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) {
        fcpt_destroy(c);
        return FCPT_ERR_DEVICE;
    }
    *out = c;
    return FCPT_OK;
}

int fcpt_destroy(fcpt_ctx *ctx)
{
    if (!ctx) return FCPT_OK;
    hipSetDevice(ctx->device);
    if (ctx->stream) {
        hipStreamSynchronize(ctx->stream);
        hipStreamDestroy(ctx->stream);
    }
    if (ctx->step_graph) hipGraphExecDestroy(ctx->step_graph);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    for (void *p : ctx->allocs) hipFree(p);
    if (ctx->s_host) hipHostFree(ctx->s_host);
    if (ctx->totals_host) hipHostFree(ctx->totals_host);
    if (ctx->force_host) hipHostFree(ctx->force_host);
    delete ctx;
    return FCPT_OK;
}

int fcpt_synchronize(fcpt_ctx *ctx)
{
    CTX_CHECK(ctx);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return FCPT_OK;
}

// read back the device scalars; flags a non-finite CFL result
static int read_scalars(fcpt_ctx *ctx)
{
    HIP_TRY(ctx, hipMemcpyAsync(ctx->s_host, ctx->s_dev, sizeof(DevScalars), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return FCPT_OK;
}

// the non-finite-dt flag is reported once: a call that reports it clears it
// on the device, as do a new clock and a new state
static int clear_error_flags(fcpt_ctx *ctx)
{
    HIP_TRY(ctx, hipMemsetAsync(&ctx->s_dev->error_flags, 0, sizeof(int), ctx->stream));
    return FCPT_OK;
}
static bool take_nonfinite(fcpt_ctx *ctx)
{
    if (!(ctx->s_host->error_flags & 1)) return false;
    ctx->s_host->error_flags &= ~1;
    clear_error_flags(ctx);
    return true;
}

// split storage: the field's background becomes its ring means, the device
// holds (float)(value - background)
static int upload_split(fcpt_ctx *ctx, int field, const double *host)
{
    const int ns = ctx->P.ns;
    const size_t rows = field_rows(field, ctx->P.nr);
    std::vector<double> &bg = ctx->hbg[field];
    bg.assign((size_t)ctx->P.nr + 2, 0.0);
    std::vector<float> d(rows * ns);
    for (size_t i = 0; i < rows; ++i) {
        const double *row = host + i * ns;
        double sum = 0.0;
        for (int j = 0; j < ns; ++j) sum += row[j];
        bg[i] = sum / (double)ns;
        for (int j = 0; j < ns; ++j) d[i * ns + j] = (float)(row[j] - bg[i]);
    }
    HIP_TRY(ctx, hipMemcpyAsync(ctx->P.bg[field], bg.data(), bg.size() * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->P.f[field], d.data(), d.size() * sizeof(float),
                                hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return FCPT_OK;
}

static int download_split(fcpt_ctx *ctx, int field, double *host)
{
    const int ns = ctx->P.ns;
    const size_t rows = field_rows(field, ctx->P.nr);
    std::vector<double> bg(rows);
    std::vector<float> d(rows * ns);
    HIP_TRY(ctx, hipMemcpyAsync(bg.data(), ctx->P.bg[field], rows * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(d.data(), ctx->P.f[field], d.size() * sizeof(float),
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < rows; ++i)
        for (int j = 0; j < ns; ++j) host[i * ns + j] = bg[i] + (double)d[i * ns + j];
    return FCPT_OK;
}

int fcpt_upload_field(fcpt_ctx *ctx, int32_t field, const double *host)
{
    CTX_SYNCED(ctx);
    if (field < 0 || field >= FCPT_NUM_FIELDS || !host) return fail(ctx, FCPT_ERR_ARG, "bad field");
    if (ctx->P.split && field != FCPT_POTENTIAL) return upload_split(ctx, field, host);
    const size_t n = field_rows(field, ctx->P.nr) * ctx->P.ns;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->P.f[field], host, n * sizeof(double), hipMemcpyHostToDevice,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return FCPT_OK;
}

int fcpt_download_field(fcpt_ctx *ctx, int32_t field, double *host)
{
    CTX_SYNCED(ctx);
    if (field < 0 || field >= FCPT_NUM_FIELDS || !host) return fail(ctx, FCPT_ERR_ARG, "bad field");
    if (ctx->P.split && field != FCPT_POTENTIAL) return download_split(ctx, field, host);
    const size_t n = field_rows(field, ctx->P.nr) * ctx->P.ns;
    HIP_TRY(ctx, hipMemcpyAsync(host, ctx->P.f[field], n * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return FCPT_OK;
}

// split storage: re-express a field about a new host-given background
static int rebase_field(fcpt_ctx *ctx, int field, const std::vector<double> &newbg)
{
    const KParams &P = ctx->P;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->bg_tmp, newbg.data(), ((size_t)P.nr + 2) * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream));
    launch_rebase(P, ctx->stream, P.f[field], P.bg[field], ctx->bg_tmp, (int)field_rows(field, P.nr));
    LAUNCH_CHECK(ctx);
    HIP_TRY(ctx, hipMemcpyAsync(P.bg[field], ctx->bg_tmp, ((size_t)P.nr + 2) * sizeof(double),
                                hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->hbg[field] = newbg;
    return FCPT_OK;
}

// split storage: backgrounds of the derived fields and of the transport
// momenta, evaluated from the ring backgrounds of Sigma, v_r, v_phi and E with
// the expressions of the kernels that produce those fields (azimuthal
// differences vanish on a ring profile)
static int set_derived_backgrounds(fcpt_ctx *ctx)
{
    const KParams &P = ctx->P;
    const int nr = P.nr;
    const std::vector<double> &S = ctx->hbg[FCPT_SIGMA], &VR = ctx->hbg[FCPT_VRAD],
                              &VA = ctx->hbg[FCPT_VAZI], &E = ctx->hbg[FCPT_ENERGY];
    const size_t n1 = (size_t)nr + 2;
    std::vector<double> cs(n1, 0.0), H(n1, 0.0), Pr(n1, 0.0), nu(n1, 0.0), T(n1, 0.0), dv(n1, 0.0),
        trr(n1, 0.0), tpp(n1, 0.0), trp(n1, 0.0), mr0(n1, 0.0), mr1(n1, 0.0), ma(n1, 0.0);
    const double g = P.gamma;
    for (int i = 0; i < nr; ++i) {
        const double R = ctx->Rmed[i];
        const double omk = std::sqrt(P.G * P.mstar / (R * R * R));
        double c, h;
        if (P.adiabatic) {
            c = std::sqrt(g * (g - 1.0) * E[i] / S[i]);
            h = c / std::sqrt(g) / omk;
            Pr[i] = (g - 1.0) * E[i];
            T[i] = P.mu / P.Rgas * (g - 1.0) * E[i] / S[i];
        } else {
            c = P.aspect_ratio * std::pow(R, P.flaring) * std::sqrt(P.G * P.mstar / R);
            h = c / omk;
            Pr[i] = S[i] * (c * c);
            T[i] = P.mu / P.Rgas * Pr[i] / S[i];
        }
        cs[i] = c;
        H[i] = h;
        nu[i] = P.alpha * h * c;
        dv[i] = (VR[i + 1] * ctx->Rinf[i + 1] - VR[i] * ctx->Rinf[i]) * ctx->InvDiffRsupRb[i];
        trr[i] = 2.0 * nu[i] * S[i] * ((VR[i + 1] - VR[i]) * ctx->InvDiffRsup[i] - dv[i] / 3.0);
        tpp[i] = 2.0 * nu[i] * S[i] * (0.5 * (VR[i + 1] + VR[i]) * ctx->InvRmed[i] - dv[i] / 3.0);
        if (i >= 1)
            trp[i] = 0.5 * (nu[i] + nu[i - 1]) * 0.5 * (S[i] + S[i - 1]) * ctx->Rinf[i] *
                     (VA[i] * ctx->InvRmed[i] - VA[i - 1] * ctx->InvRmed[i - 1]) * ctx->InvDiffRmed[i];
        mr0[i] = S[i] * VR[i + 1];
        mr1[i] = S[i] * VR[i];
        ma[i] = S[i] * (VA[i] + R * P.omega_frame) * R;
    }
    if (!std::isfinite(cs[1]) || !std::isfinite(trp[1]))
        return fail(ctx, FCPT_ERR_ARG, "split storage needs a positive Sigma (and E) on every ring");
    const std::pair<int, const std::vector<double> *> derived[] = {
        {FCPT_SOUNDSPEED, &cs}, {FCPT_SCALE_HEIGHT, &H},  {FCPT_PRESSURE, &Pr},
        {FCPT_VISCOSITY, &nu},  {FCPT_TEMPERATURE, &T},   {FCPT_DIV_V, &dv},
        {FCPT_TAU_R_R, &trr},   {FCPT_TAU_PHI_PHI, &tpp}, {FCPT_TAU_R_PHI, &trp}};
    for (const auto &fd : derived) {
        const int rc = rebase_field(ctx, fd.first, *fd.second);
        if (rc) return rc;
    }
    const std::vector<double> *mom[4] = {&mr0, &mr1, &ma, &ma};
    for (int k = 0; k < 4; ++k)
        HIP_TRY(ctx, hipMemcpyAsync(P.bgm[k], mom[k]->data(), n1 * sizeof(double),
                                    hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return FCPT_OK;
}

// split storage: the backgrounds of the evolved fields move to their current
// ring means (each deviation re-expressed about the new background, rounded
// to nearest once) and the derived fields' and momenta's backgrounds follow
static int rebase_backgrounds(fcpt_ctx *ctx)
{
    const KParams &P = ctx->P;
    ctx->steps_since_rebase = 0;
    if (!P.split) return FCPT_OK;
    int rc = sync_derived(ctx);
    if (rc) return rc;
    const int state[4] = {FCPT_SIGMA, FCPT_VRAD, FCPT_VAZI, FCPT_ENERGY};
    for (int k = 0; k < (P.adiabatic ? 4 : 3); ++k) {
        const int f = state[k];
        const int rows = (int)field_rows(f, P.nr);
        std::vector<double> bg((size_t)P.nr + 2, 0.0);
        launch_split_ring_mean(P, ctx->stream, P.f[f], P.bg[f], rows, ctx->bg_tmp);
        LAUNCH_CHECK(ctx);
        HIP_TRY(ctx, hipMemcpyAsync(bg.data(), ctx->bg_tmp, (size_t)rows * sizeof(double),
                                    hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if ((rc = rebase_field(ctx, f, bg))) return rc;
    }
    return set_derived_backgrounds(ctx);
}

int fcpt_rebase_backgrounds(fcpt_ctx *ctx)
{
    CTX_CHECK(ctx);
    return rebase_backgrounds(ctx);
}

// copy_initial_values (damping.cpp:301-309): the damping references take the
// current fields (split storage: with their backgrounds)
static int copy_initial_values(fcpt_ctx *ctx)
{
    const KParams &P = ctx->P;
    const int pairs[4][2] = {{FCPT_SIGMA, FCPT_SIGMA0}, {FCPT_VRAD, FCPT_VRAD0},
                             {FCPT_VAZI, FCPT_VAZI0}, {FCPT_ENERGY, FCPT_ENERGY0}};
    for (const auto &pr : pairs) {
        const size_t bytes = field_rows(pr[0], P.nr) * P.ns * element_size(P);
        HIP_TRY(ctx, hipMemcpyAsync(P.f[pr[1]], P.f[pr[0]], bytes, hipMemcpyDeviceToDevice, ctx->stream));
        if (P.split) {
            ctx->hbg[pr[1]] = ctx->hbg[pr[0]];
            HIP_TRY(ctx, hipMemcpyAsync(P.bg[pr[1]], P.bg[pr[0]], ((size_t)P.nr + 2) * sizeof(double),
                                        hipMemcpyDeviceToDevice, ctx->stream));
        }
    }
    return FCPT_OK;
}

int fcpt_init_profiles(fcpt_ctx *ctx, const fcpt_ic_desc *ic)
{
    CTX_SYNCED(ctx);
    if (!ic) return fail(ctx, FCPT_ERR_ARG, "null initial-condition description");
    const KParams &P = ctx->P;
    const int nr = P.nr, ns = P.ns;
    const DiskProfile prof{ctx->desc, *ic};
    // ring profiles, O(nr) on the host
    std::vector<double> sig((size_t)nr + 2, 0.0), en((size_t)nr + 2, 0.0), vr((size_t)nr + 2, 0.0),
        vaz((size_t)nr + 2, 0.0), zero((size_t)nr + 2, 0.0);
    for (int i = 0; i < nr; ++i) {
        sig[i] = prof.sigma_cell(ctx->Rmed[i]);
        if (P.adiabatic) en[i] = DiskProfile::max_(prof.energy(ctx->Rmed[i]), prof.energy_floor(sig[i]));
    }
    // SigmaMed and SigmaInf (compute_azi_avg_Sigma, Theo.cpp:30-48): the ring
    // sum of nphi equal values, in order, divided by nphi
    std::vector<double> smed((size_t)nr, 0.0), sinf((size_t)nr, 0.0);
    for (int i = 0; i < nr; ++i) {
        double m = 0.0;
        for (int j = 0; j < ns; ++j) m += sig[i];
        smed[i] = m / (double)ns;
    }
    sinf[0] = smed[0];
    for (int i = 1; i < nr; ++i) {
        const double dr = ctx->Rmed[i] - ctx->Rmed[i - 1];
        sinf[i] = (smed[i - 1] * (ctx->Rmed[i] - ctx->Rinf[i]) + smed[i] * (ctx->Rinf[i] - ctx->Rmed[i - 1])) / dr;
    }
    for (int i = 0; i < nr; ++i) {
        const double r = ctx->Rmed[i], ri = ctx->Rinf[i];
        double v = prof.vphi(r);
        v -= P.omega_frame * r;
        vaz[i] = v;
        double u = ic->imposed_disk_drift * P.sigma0 / sinf[i] / ri;
        if (!ic->vradial_zero) u += prof.vr_viscous(ri);
        else u = 0.0;
        vr[i] = u;
    }
    // the rings on the device (one allocation per call, freed at the end)
    double *d_rings = nullptr;
    const size_t n1 = (size_t)nr + 2;
    HIP_TRY(ctx, hipMalloc(&d_rings, 5 * n1 * sizeof(double)));
    double *d_sig = d_rings, *d_en = d_rings + n1, *d_vr = d_rings + 2 * n1, *d_vaz = d_rings + 3 * n1,
           *d_zero = d_rings + 4 * n1;
    int rc = FCPT_OK;
    auto up = [&](double *dst, const std::vector<double> &src) {
        if (!rc && hipMemcpyAsync(dst, src.data(), n1 * sizeof(double), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            rc = fail(ctx, FCPT_ERR_DEVICE, "ring profile upload");
    };
    up(d_sig, sig);
    up(d_en, en);
    up(d_vr, vr);
    up(d_vaz, vaz);
    up(d_zero, zero);
    // split storage: each field's background is its ring profile, so every
    // deviation is zero (the velocities' background is zero while init_euler
    // runs, then their profiles)
    auto fill = [&](int f, const std::vector<double> &ring, const double *d_ring, int rows) {
        const double *b = nullptr;
        if (P.split) {
            ctx->hbg[f] = ring;
            if (!rc && hipMemcpyAsync(P.bg[f], ring.data(), n1 * sizeof(double), hipMemcpyHostToDevice,
                                      ctx->stream) != hipSuccess)
                rc = fail(ctx, FCPT_ERR_DEVICE, "background upload");
            b = P.bg[f];
        }
        if (!rc) launch_fill_rows(P, ctx->stream, P.f[f], d_ring, b, rows);
    };
    fill(FCPT_SIGMA, sig, d_sig, nr);
    fill(FCPT_ENERGY, en, d_en, nr);
    // init_euler runs while the velocities are still zero (init.cpp:331-332)
    fill(FCPT_VRAD, zero, d_zero, nr + 1);
    fill(FCPT_VAZI, zero, d_zero, nr);
    if (!rc && P.split) rc = set_derived_backgrounds(ctx);
    if (!rc) {
        if (!P.adiabatic) {
            launch_iso_profile(P, ctx->stream);
            launch_recalc_viscosity(P, ctx->stream);
        } else {
            launch_derived(P, ctx->stream);
            launch_recalc_viscosity(P, ctx->stream);
            launch_stress(P, ctx->stream);
            launch_heatcool_cfl(P, ctx->stream);
        }
    }
    fill(FCPT_VRAD, vr, d_vr, nr + 1);
    fill(FCPT_VAZI, vaz, d_vaz, nr);
    if (!rc && P.split) rc = set_derived_backgrounds(ctx);
    if (!rc) rc = copy_initial_values(ctx);
    if (!rc) launch_boundaries(P, ctx->stream, 0, 0);
    if (!rc) rc = copy_initial_values(ctx);
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipStreamSynchronize(ctx->stream);
    hipFree(d_rings);
    if (rc) return rc;
    if (le != hipSuccess) return fail(ctx, FCPT_ERR_DEVICE, "launch: %s", hipGetErrorString(le));
    if (se != hipSuccess) return fail(ctx, FCPT_ERR_DEVICE, "sync: %s", hipGetErrorString(se));
    return FCPT_OK;
}

int fcpt_upload_reference(fcpt_ctx *ctx, const double *sigma0, const double *vrad0,
                          const double *vazi0, const double *energy0)
{
    int rc;
    if ((rc = fcpt_upload_field(ctx, FCPT_SIGMA0, sigma0))) return rc;
    if ((rc = fcpt_upload_field(ctx, FCPT_VRAD0, vrad0))) return rc;
    if ((rc = fcpt_upload_field(ctx, FCPT_VAZI0, vazi0))) return rc;
    if (energy0) return fcpt_upload_field(ctx, FCPT_ENERGY0, energy0);
    return FCPT_OK;
}

int fcpt_upload_state(fcpt_ctx *ctx, const double *sigma, const double *vrad, const double *vazi,
                      const double *energy)
{
    int rc;
    if ((rc = fcpt_upload_field(ctx, FCPT_SIGMA, sigma))) return rc;
    if ((rc = fcpt_upload_field(ctx, FCPT_VRAD, vrad))) return rc;
    if ((rc = fcpt_upload_field(ctx, FCPT_VAZI, vazi))) return rc;
    if (energy && (rc = fcpt_upload_field(ctx, FCPT_ENERGY, energy))) return rc;
    if (ctx->P.adiabatic && !energy) return fail(ctx, FCPT_ERR_ARG, "adiabatic disk needs energy");
    if (ctx->P.split && (rc = set_derived_backgrounds(ctx))) return rc;
    // init_euler recalculation (SourceEuler.cpp:264-284)
    const KParams &P = ctx->P;
    if (!P.adiabatic) {
        launch_iso_profile(P, ctx->stream);
        launch_recalc_viscosity(P, ctx->stream);
    } else {
        launch_derived(P, ctx->stream);
        // compute_heating_cooling_for_CFL (SourceEuler.cpp:1410-1450)
        launch_recalc_viscosity(P, ctx->stream);
        launch_stress(P, ctx->stream);
        launch_heatcool_cfl(P, ctx->stream);
    }
    LAUNCH_CHECK(ctx);
    if ((rc = clear_error_flags(ctx))) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return FCPT_OK;
}

int fcpt_download_state(fcpt_ctx *ctx, double *sigma, double *vrad, double *vazi, double *energy)
{
    int rc;
    if (sigma && (rc = fcpt_download_field(ctx, FCPT_SIGMA, sigma))) return rc;
    if (vrad && (rc = fcpt_download_field(ctx, FCPT_VRAD, vrad))) return rc;
    if (vazi && (rc = fcpt_download_field(ctx, FCPT_VAZI, vazi))) return rc;
    if (energy && (rc = fcpt_download_field(ctx, FCPT_ENERGY, energy))) return rc;
    return FCPT_OK;
}

int fcpt_set_time(fcpt_ctx *ctx, double time, double last_dt, int32_t n_monitor)
{
    CTX_CHECK(ctx);
    launch_set_time(ctx->P, ctx->stream, time, last_dt, n_monitor);
    LAUNCH_CHECK(ctx);
    return clear_error_flags(ctx);
}

int fcpt_get_time(fcpt_ctx *ctx, double *time, double *last_dt, int32_t *n_monitor)
{
    CTX_CHECK(ctx);
    int rc = read_scalars(ctx);
    if (rc) return rc;
    if (time) *time = ctx->s_host->time;
    if (last_dt) *last_dt = ctx->s_host->last_dt;
    if (n_monitor) *n_monitor = ctx->s_host->n_monitor;
    return FCPT_OK;
}

// compute_smoothing (Force.cpp:145-159) per body, on the host with the
// reference's expressions (compute_smoothing_iso_planet :132-142, the body's
// r = hypot(x, y) as planet.cpp:159 keeps it); < 0 marks the default,
// ThicknessSmoothing times the scale height of the cell
static BodySmoothing body_smoothing(const fcpt_ctx *ctx, const fcpt_step_params &sp)
{
    BodySmoothing b{};
    const fcpt_disk_desc &d = ctx->desc;
    for (int k = 0; k < kMaxPlanets; ++k) {
        b.s[k] = -1.0;
        if (k >= sp.n_planets) continue;
        if ((d.potential_smoothing & FCPT_SMOOTHING_NO_STAR) && k == 0) {
            b.s[k] = 0;
        } else if (d.potential_smoothing & FCPT_SMOOTHING_PLANETLOC) {
            const double a = std::hypot(sp.planet[k].x, sp.planet[k].y);
            const double scale_height = d.aspect_ratio * std::pow(a, 1 + d.flaring_index);
            b.s[k] = d.thickness_smoothing * scale_height;
        }
    }
    return b;
}

int fcpt_set_step_params(fcpt_ctx *ctx, const fcpt_step_params *params)
{
    CTX_CHECK(ctx);
    if (!params || params->n_planets < 0 || params->n_planets > FCPT_MAX_PLANETS)
        return fail(ctx, FCPT_ERR_ARG, "bad step parameters");
    launch_set_params(ctx->P, ctx->stream, *params, body_smoothing(ctx, *params));
    LAUNCH_CHECK(ctx);
    return FCPT_OK;
}

static int cfl_common(fcpt_ctx *ctx, int mode, double *dt_out)
{
    CTX_SYNCED(ctx);
    launch_cfl(ctx->P, ctx->stream, mode);
    LAUNCH_CHECK(ctx);
    int rc = read_scalars(ctx);
    if (rc) return rc;
    if (take_nonfinite(ctx))
        return fail(ctx, FCPT_ERR_NONFINITE, "CFL time step is not finite and positive (%g)",
                    ctx->s_host->cfl_dt);
    if (dt_out) *dt_out = ctx->s_host->cfl_dt;
    return FCPT_OK;
}

int fcpt_condition_cfl(fcpt_ctx *ctx, double *dt_out) { return cfl_common(ctx, 0, dt_out); }
int fcpt_calculate_timestep(fcpt_ctx *ctx, double *dt_out) { return cfl_common(ctx, 1, dt_out); }

int fcpt_nbody_potential(fcpt_ctx *ctx)
{
    CTX_SYNCED(ctx);
    launch_potential(ctx->P, ctx->stream);
    LAUNCH_CHECK(ctx);
    return FCPT_OK;
}

#define DT_ENTRY(name, body)                         \
    int name(fcpt_ctx *ctx, double dt)               \
    {                                                \
        CTX_SYNCED(ctx);                             \
        launch_set_dt(ctx->P, ctx->stream, dt);      \
        body;                                        \
        LAUNCH_CHECK(ctx);                           \
        return FCPT_OK;                              \
    }

DT_ENTRY(fcpt_update_with_sourceterms, launch_sources(ctx->P, ctx->stream))
DT_ENTRY(fcpt_artificial_viscosity, launch_artificial_viscosity(ctx->P, ctx->stream, 0))
DT_ENTRY(fcpt_update_velocities_with_viscosity, launch_visc_update(ctx->P, ctx->stream))
DT_ENTRY(fcpt_transport, launch_transport(ctx->P, ctx->stream))

int fcpt_recalculate_viscosity(fcpt_ctx *ctx)
{
    CTX_SYNCED(ctx);
    launch_recalc_viscosity(ctx->P, ctx->stream);
    LAUNCH_CHECK(ctx);
    return FCPT_OK;
}

int fcpt_viscous_stress_tensor(fcpt_ctx *ctx)
{
    CTX_SYNCED(ctx);
    launch_stress(ctx->P, ctx->stream);
    LAUNCH_CHECK(ctx);
    return FCPT_OK;
}

int fcpt_substep3(fcpt_ctx *ctx, double time, double dt)
{
    CTX_SYNCED(ctx);
    if (!ctx->P.adiabatic) return FCPT_OK;
    launch_set_clock(ctx->P, ctx->stream, time);
    launch_set_dt(ctx->P, ctx->stream, dt);
    launch_substep3(ctx->P, ctx->stream);
    LAUNCH_CHECK(ctx);
    return FCPT_OK;
}

int fcpt_apply_boundary_condition(fcpt_ctx *ctx, double dt, int32_t final_call)
{
    CTX_SYNCED(ctx);
    launch_set_dt(ctx->P, ctx->stream, dt);
    launch_boundaries(ctx->P, ctx->stream, final_call, 0);
    LAUNCH_CHECK(ctx);
    return FCPT_OK;
}

int fcpt_recalculate_derived(fcpt_ctx *ctx)
{
    CTX_SYNCED(ctx);
    launch_derived(ctx->P, ctx->stream);
    LAUNCH_CHECK(ctx);
    return FCPT_OK;
}

int fcpt_get_nshift(fcpt_ctx *ctx, int32_t *nshift)
{
    CTX_CHECK(ctx);
    if (!nshift) return fail(ctx, FCPT_ERR_ARG, "null output");
    HIP_TRY(ctx, hipMemcpyAsync(nshift, ctx->P.nshift, sizeof(int32_t) * ctx->P.nr,
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return FCPT_OK;
}

// the gas part of step_Euler (simulation.cpp:148-267); dt already in device memory
// The source part of a step (potential, sources, artificial viscosity,
// viscosity recalculation, viscous stress and update, SubStep3) with the dt
// and time in device memory.  The fused form runs the source terms, the
// artificial viscosity and the viscous stress + update as one pass each; each
// pass writes the velocities to the partner buffers, so the passes alternate
// between the two views of the state (P and its swapped view).  Returns the
// view that holds the velocities afterwards.  fold_mode >= 0 (fused only):
// the potential launch also folds the CFL minima of launch_cfl_cells.
static const KParams *enqueue_kick(const KParams &P, const KParams &Q, hipStream_t st, bool fused,
                                   bool pressure_first, int fold_mode = -1)
{
    if (fused && kick_walk_applies(P)) {
        // the potential pass and two walks: sources and artificial
        // viscosity; then the viscous stress and update (and SubStep3)
        if (pressure_first && P.adiabatic) launch_pressure(P, st);
        launch_kick(P, st, fold_mode);
        if (P.adiabatic) launch_visc_energy(Q, st);
        else launch_visc_fused(Q, st);
        return &P;
    }
    if (fused) launch_potential_fold(P, st, fold_mode);
    else launch_potential(P, st);
    if (!fused) {
        if (pressure_first) launch_pressure(P, st);
        launch_sources(P, st);
        if (!launch_artificial_viscosity(P, st, 1)) launch_recalc_viscosity(P, st);
        launch_stress(P, st);
        launch_visc_update(P, st);
        if (P.adiabatic) launch_substep3(P, st);
        return &P;
    }
    const KParams *cur = &P, *other = &Q;
    auto flip = [&]() { std::swap(cur, other); };
    // the isothermal fused source pass forms the pressure from Sigma itself
    if (pressure_first && P.adiabatic) launch_pressure(P, st);
    launch_src_fused(*cur, st);
    flip();
    if (P.adiabatic) launch_src_energy(*cur, st);
    if (P.artvisc == FCPT_ARTVISC_TW || P.artvisc == FCPT_ARTVISC_SN) {
        launch_av_fused(*cur, st, 1);
        flip();
    } else {
        launch_artificial_viscosity(*cur, st, 1);  // temperature floor only
        launch_recalc_viscosity(*cur, st);
    }
    if (P.adiabatic) launch_visc_energy(*cur, st);
    else launch_visc_fused(*cur, st);
    flip();
    return cur;
}

// One gas step of step_Euler; the transport's velocity kernel writes the
// final velocities to the fields of view P and, in the fused step, applies
// the damping of the final boundary call that follows it.
static void enqueue_euler(const KParams &P, hipStream_t st, int count_monitor, bool fused, int fold_mode)
{
    const KParams Q = swapped_view(P);
    const KParams *cur = enqueue_kick(P, Q, st, fused, false, fused ? fold_mode : -1);
    if (fused) {
        // the transport's velocity kernel also runs the final boundary call
        // and the derived fields
        launch_transport_ghost(*cur, P, st, count_monitor ? 2 : 1);
    } else {
        launch_boundaries(*cur, st, 0, 0);
        launch_transport_to(*cur, P, st);
        launch_final_boundaries_derived(P, st, count_monitor ? 2 : 1);
    }
}

// One gas step of step_LeapFrog (simulation.cpp:276-459): kick with dt / 2,
// boundaries and transport over dt, kick with dt / 2 at the midstep time and
// bodies (compute_pressure first), final boundaries over dt, derived fields.
// The isothermal unfused path needs the pressure of the transported Sigma,
// as does the adiabatic one of the transported E.
static void enqueue_leapfrog(const KParams &P, hipStream_t st, int count_monitor, bool fused)
{
    const KParams Q = swapped_view(P);
    launch_leapfrog_phase(P, st, 0);
    const KParams *cur = enqueue_kick(P, Q, st, fused, false);
    launch_leapfrog_phase(P, st, 1);
    if (fused) {
        launch_transport_ghost(*cur, P, st, 0);
    } else {
        launch_boundaries(*cur, st, 0, 0);
        launch_transport_to(*cur, P, st);
    }
    launch_leapfrog_phase(P, st, 2);
    cur = enqueue_kick(P, Q, st, fused, true);
    if (cur != &P) {
        // an odd number of fused passes: the velocities go back to the fields
        launch_copy_velocities(*cur, st);
    }
    launch_leapfrog_phase(P, st, 3);
    launch_final_boundaries_derived(P, st, count_monitor ? 2 : 1);
}

static void enqueue_step(const KParams &P, hipStream_t st, int count_monitor, bool fused)
{
    if (P.integrator == FCPT_INTEGRATOR_LEAPFROG) enqueue_leapfrog(P, st, count_monitor, fused);
    else enqueue_euler(P, st, count_monitor, fused, -1);
}

// one sim::run iteration: CalculateTimeStep with the monitor clip (mode 2),
// then the step; the fused Euler step folds the CFL minima inside its
// potential launch (one launch less)
static void enqueue_iteration(const fcpt_ctx *ctx)
{
    const KParams &P = ctx->P;
    if (ctx->fused && P.integrator != FCPT_INTEGRATOR_LEAPFROG) {
        const KParams V = step_view(ctx);
        launch_cfl_cells(V, ctx->stream);
        enqueue_euler(V, ctx->stream, 1, true, 2);
    } else {
        launch_cfl(P, ctx->stream, 2);
        enqueue_step(P, ctx->stream, 1, ctx->fused);
    }
}

int fcpt_step_euler(fcpt_ctx *ctx, double dt)
{
    CTX_SYNCED(ctx);
    launch_set_dt(ctx->P, ctx->stream, dt);
    enqueue_euler(step_view(ctx), ctx->stream, 0, ctx->fused, -1);
    LAUNCH_CHECK(ctx);
    ctx->derived_stale = lazy_applies(ctx);
    return FCPT_OK;
}

int fcpt_step_leapfrog(fcpt_ctx *ctx, double dt)
{
    CTX_SYNCED(ctx);
    launch_set_dt(ctx->P, ctx->stream, dt);
    enqueue_leapfrog(ctx->P, ctx->stream, 0, ctx->fused);
    LAUNCH_CHECK(ctx);
    return FCPT_OK;
}

int fcpt_compute_pressure(fcpt_ctx *ctx)
{
    CTX_SYNCED(ctx);
    launch_pressure(ctx->P, ctx->stream);
    LAUNCH_CHECK(ctx);
    return FCPT_OK;
}

int fcpt_set_midstep_params(fcpt_ctx *ctx, const fcpt_step_params *params)
{
    CTX_CHECK(ctx);
    if (!params || params->n_planets < 0 || params->n_planets > FCPT_MAX_PLANETS)
        return fail(ctx, FCPT_ERR_ARG, "bad step parameters");
    launch_set_mid_params(ctx->P, ctx->stream, *params, body_smoothing(ctx, *params));
    LAUNCH_CHECK(ctx);
    return FCPT_OK;
}

int fcpt_run_step(fcpt_ctx *ctx, double *dt_out)
{
    CTX_CHECK(ctx);
    enqueue_iteration(ctx);
    LAUNCH_CHECK(ctx);
    if (lazy_applies(ctx)) ctx->derived_stale = true;
    int rc = read_scalars(ctx);
    if (rc) return rc;
    if (take_nonfinite(ctx))
        return fail(ctx, FCPT_ERR_NONFINITE, "CFL time step is not finite and positive");
    if (dt_out) *dt_out = ctx->s_host->dt;
    return FCPT_OK;
}

// capture one sim::run iteration into a graph (once per context)
static int ensure_step_graph(fcpt_ctx *ctx)
{
    if (ctx->step_graph) return FCPT_OK;
    hipGraph_t g = nullptr;
    HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    enqueue_iteration(ctx);
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (e != hipSuccess) return fail(ctx, FCPT_ERR_DEVICE, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&ctx->step_graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        ctx->step_graph = nullptr;
        return fail(ctx, FCPT_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(e));
    }
    return FCPT_OK;
}

// split storage with a rebase interval: the steps before the next background move
static bool periodic_rebase(const fcpt_ctx *ctx) { return ctx->P.split && ctx->rebase_interval > 0; }

int fcpt_run_steps(fcpt_ctx *ctx, int32_t n_steps, int32_t use_graph)
{
    CTX_CHECK(ctx);
    if (n_steps < 0) return fail(ctx, FCPT_ERR_ARG, "negative step count");
    if (use_graph) {
        int rc = ensure_step_graph(ctx);
        if (rc) return rc;
    }
    for (int left = n_steps; left > 0;) {
        int m = left;
        if (periodic_rebase(ctx)) m = std::min(m, ctx->rebase_interval - ctx->steps_since_rebase);
        if (use_graph) {
            for (int k = 0; k < m; ++k) HIP_TRY(ctx, hipGraphLaunch(ctx->step_graph, ctx->stream));
        } else {
            for (int k = 0; k < m; ++k) enqueue_iteration(ctx);
            LAUNCH_CHECK(ctx);
        }
        left -= m;
        if (lazy_applies(ctx)) ctx->derived_stale = true;
        if (periodic_rebase(ctx) && (ctx->steps_since_rebase += m) >= ctx->rebase_interval) {
            const int rc = rebase_backgrounds(ctx);
            if (rc) return rc;
        }
    }
    return FCPT_OK;
}

static int ensure_events(fcpt_ctx *ctx)
{
    if (!ctx->ev0) HIP_TRY(ctx, hipEventCreate(&ctx->ev0));
    if (!ctx->ev1) HIP_TRY(ctx, hipEventCreate(&ctx->ev1));
    return FCPT_OK;
}

int fcpt_time_steps(fcpt_ctx *ctx, int32_t n_steps, int32_t use_graph, double *ms_total)
{
    CTX_CHECK(ctx);
    int rc = ensure_events(ctx);
    if (!rc && use_graph) rc = ensure_step_graph(ctx);
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    rc = fcpt_run_steps(ctx, n_steps, use_graph);
    if (rc) return rc;
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    rc = read_scalars(ctx);
    if (rc) return rc;
    if (take_nonfinite(ctx))
        return fail(ctx, FCPT_ERR_NONFINITE, "CFL time step became non-finite during the run");
    if (ms_total) *ms_total = (double)ms;
    return FCPT_OK;
}

int fcpt_time_kernel(fcpt_ctx *ctx, int32_t kernel, int32_t reps, double *ms_per_launch)
{
    CTX_SYNCED(ctx);
    if (reps <= 0) return fail(ctx, FCPT_ERR_ARG, "reps must be positive");
    int rc = ensure_events(ctx);
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ForceBodies fb{};
    fb.n = 2;
    fb.flags = FCPT_FORCE_CORRECT_SELFGRAVITY;
    fb.x[1] = fb.r[1] = 1.0;
    fb.smooth_fixed[0] = fb.smooth_fixed[1] = -1.0;
    auto one = [&]() {
        switch (kernel) {
        case FCPT_KERNEL_DISK_FORCE:
            launch_disk_force(ctx->P, ctx->stream, fb, ctx->force_sig1d, ctx->force_part,
                              ctx->force_out);
            break;
        case FCPT_KERNEL_RADIAL_SWEEP: launch_radial_sweep(ctx->P, ctx->stream); break;
        case FCPT_KERNEL_AZIMUTHAL: launch_azimuthal_sweep(ctx->P, ctx->stream); break;
        case FCPT_KERNEL_CFL: launch_cfl_cells(ctx->P, ctx->stream); break;
        case FCPT_KERNEL_STRESS: launch_stress(ctx->P, ctx->stream); break;
        case FCPT_KERNEL_POTENTIAL: launch_potential(ctx->P, ctx->stream); break;
        case FCPT_KERNEL_VISC_FUSED: launch_visc_fused(ctx->P, ctx->stream); break;
        default: break;
        }
    };
    if (kernel < 0 || kernel > FCPT_KERNEL_VISC_FUSED || kernel == 2)
        return fail(ctx, FCPT_ERR_ARG, "unknown kernel");
    one();  // warm-up
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int k = 0; k < reps; ++k) one();
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    LAUNCH_CHECK(ctx);
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (ms_per_launch) *ms_per_launch = (double)ms / reps;
    return FCPT_OK;
}

int fcpt_time_copy(fcpt_ctx *ctx, int32_t reps, double *ms_per_launch, double *bytes_per_launch)
{
    CTX_CHECK(ctx);
    if (reps <= 0) return fail(ctx, FCPT_ERR_ARG, "reps must be positive");
    int rc = ensure_events(ctx);
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    launch_copy4(ctx->P, ctx->stream);
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int k = 0; k < reps; ++k) launch_copy4(ctx->P, ctx->stream);
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    LAUNCH_CHECK(ctx);
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (ms_per_launch) *ms_per_launch = (double)ms / reps;
    if (bytes_per_launch)
        *bytes_per_launch = 2.0 * 4.0 * (double)((size_t)ctx->P.nr * ctx->P.ns * element_size(ctx->P) / 8 * 8);
    return FCPT_OK;
}

int fcpt_set_launch_param(fcpt_ctx *ctx, int32_t param, int32_t value)
{
    CTX_CHECK(ctx);
    if (param == FCPT_PARAM_RADIAL_TILE) {
        if (value < 0 || value > 64) return fail(ctx, FCPT_ERR_ARG, "radial tile must be 0 .. 64");
    } else if (param == FCPT_PARAM_RADIAL_WAVES) {
        if (value != 2 && value != 3) return fail(ctx, FCPT_ERR_ARG, "radial waves must be 2 or 3");
    } else if (param == FCPT_PARAM_FUSED_STEP) {
        if (value != 0 && value != 1) return fail(ctx, FCPT_ERR_ARG, "fused step must be 0 or 1");
    } else if (param == FCPT_PARAM_VISC_TILE) {
        if (value < 0 || value > 64) return fail(ctx, FCPT_ERR_ARG, "viscous tile must be 0 .. 64");
    } else if (param == FCPT_PARAM_KICK_TILE) {
        if (value < 0 || value > 64) return fail(ctx, FCPT_ERR_ARG, "kick tile must be 0 .. 64");
    } else if (param == FCPT_PARAM_SPLIT_ROUNDING) {
        if (value < 0 || value > 2) return fail(ctx, FCPT_ERR_ARG, "split rounding must be 0, 1 or 2");
    } else if (param == FCPT_PARAM_REBASE_INTERVAL) {
        if (value < 0) return fail(ctx, FCPT_ERR_ARG, "rebase interval must be 0 or positive");
        ctx->rebase_interval = ctx->P.split ? value : 0;
        ctx->steps_since_rebase = 0;
        return FCPT_OK;
    } else {
        return fail(ctx, FCPT_ERR_ARG, "unknown launch parameter");
    }
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (param == FCPT_PARAM_RADIAL_TILE) ctx->P.radial_tile = value;
    else if (param == FCPT_PARAM_RADIAL_WAVES) ctx->P.radial_waves = value;
    else if (param == FCPT_PARAM_VISC_TILE) ctx->P.visc_tile = value;
    else if (param == FCPT_PARAM_KICK_TILE) ctx->P.kick_tile = value;
    else if (param == FCPT_PARAM_SPLIT_ROUNDING) ctx->P.split_rounding = ctx->P.split ? value : 0;
    else ctx->fused = value != 0;
    if (ctx->step_graph) {
        HIP_TRY(ctx, hipGraphExecDestroy(ctx->step_graph));
        ctx->step_graph = nullptr;
    }
    return FCPT_OK;
}

// The walks' tiles by whole-step time: each candidate runs `reps` graph-replayed
// sim::run iterations from the same state, which is saved first (every field
// and the device scalars) and put back after each trial and at the end, so
// the context's state and clock come out as they went in.  Tiles 2 .. 16
// for each walk that applies, the kick walk's first, then the viscous walk's.
int fcpt_autotune_walks(fcpt_ctx *ctx, int32_t reps, int32_t *kick_tile, int32_t *visc_tile)
{
    CTX_SYNCED(ctx);
    if (reps <= 0) return fail(ctx, FCPT_ERR_ARG, "reps must be positive");
    int rc = ensure_events(ctx);
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    KParams &P = ctx->P;
    const int kick0 = resolve_kick_tile(P), visc0 = resolve_visc_tile(P);
    if (!ctx->fused || P.integrator != FCPT_INTEGRATOR_EULER || visc0 <= 1) {
        if (kick_tile) *kick_tile = kick0;
        if (visc_tile) *visc_tile = visc0;
        return FCPT_OK;  // the walks do not run in this context's step
    }
    // the trials' steps neither move the split backgrounds (the saved
    // deviations go back about the same backgrounds) nor count toward the next move
    struct RebaseHold {
        fcpt_ctx *c;
        int interval, since;
        ~RebaseHold()
        {
            c->rebase_interval = interval;
            c->steps_since_rebase = since;
        }
    } hold{ctx, ctx->rebase_interval, ctx->steps_since_rebase};
    ctx->rebase_interval = 0;
    // the saved state: every field and the scalars
    size_t total = 0;
    std::vector<size_t> off(FCPT_NUM_FIELDS);
    for (int k = 0; k < FCPT_NUM_FIELDS; ++k) {
        off[k] = total;
        total += (field_rows(k, P.nr) * P.ns * field_element_size(P, k) + 255) / 256 * 256;
    }
    char *save = nullptr;
    HIP_TRY(ctx, hipMalloc((void **)&save, total));
    DevScalars s0;
    auto copy_fields = [&](bool to_save) -> hipError_t {
        for (int k = 0; k < FCPT_NUM_FIELDS; ++k) {
            const size_t bytes = field_rows(k, P.nr) * P.ns * field_element_size(P, k);
            hipError_t e = to_save ? hipMemcpyAsync(save + off[k], P.f[k], bytes, hipMemcpyDeviceToDevice, ctx->stream)
                                   : hipMemcpyAsync(P.f[k], save + off[k], bytes, hipMemcpyDeviceToDevice, ctx->stream);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    auto restore = [&]() -> hipError_t {
        hipError_t e = copy_fields(false);
        if (e == hipSuccess) e = hipMemcpyAsync(ctx->s_dev, &s0, sizeof(DevScalars), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        ctx->derived_stale = false;
        return e;
    };
    hipError_t e = copy_fields(true);
    if (e == hipSuccess) e = hipMemcpy(&s0, ctx->s_dev, sizeof(DevScalars), hipMemcpyDeviceToHost);
    auto trial = [&](double &ms_out) -> int {
        if (ctx->step_graph) {
            hipGraphExecDestroy(ctx->step_graph);
            ctx->step_graph = nullptr;
        }
        int r = fcpt_run_steps(ctx, 1, 1);  // capture + warm-up
        if (!r) r = fcpt_time_steps(ctx, reps, 1, &ms_out);
        const hipError_t er = restore();
        if (!r && er != hipSuccess) r = fail(ctx, FCPT_ERR_DEVICE, "state restore: %s", hipGetErrorString(er));
        return r;
    };
    const int cand[] = {2, 4, 8, 16};
    int best_k = kick0, best_v = visc0;
    rc = e == hipSuccess ? FCPT_OK : fail(ctx, FCPT_ERR_DEVICE, "state save: %s", hipGetErrorString(e));
    if (!rc && kick_walk_applies(P)) {
        double best = 1e300;
        for (int t : cand) {
            if (t > P.nr) continue;
            P.kick_tile = t;
            if (!kick_walk_applies(P)) continue;
            double ms = 0.0;
            if ((rc = trial(ms))) break;
            if (ms < best) {
                best = ms;
                best_k = t;
            }
        }
        P.kick_tile = best_k;
    }
    if (!rc) {
        double best = 1e300;
        for (int t : cand) {
            if (t > P.nr) continue;
            P.visc_tile = t;
            if (P.kick_tile > 1 && !kick_walk_applies(P)) continue;
            double ms = 0.0;
            if ((rc = trial(ms))) break;
            if (ms < best) {
                best = ms;
                best_v = t;
            }
        }
        P.visc_tile = best_v;
    }
    hipFree(save);
    if (ctx->step_graph) {
        hipGraphExecDestroy(ctx->step_graph);
        ctx->step_graph = nullptr;
    }
    if (rc) return rc;
    if (kick_tile) *kick_tile = best_k;
    if (visc_tile) *visc_tile = best_v;
    return FCPT_OK;
}

int fcpt_get_launch_param(fcpt_ctx *ctx, int32_t param, int32_t *value)
{
    if (!ctx || !value) return FCPT_ERR_ARG;
    const KParams &P = ctx->P;
    switch (param) {
    case FCPT_PARAM_RADIAL_TILE: *value = radial_tile_for(P.nr, P.ns, P.radial_tile); break;
    case FCPT_PARAM_RADIAL_WAVES: *value = P.radial_waves; break;
    case FCPT_PARAM_FUSED_STEP: *value = ctx->fused ? 1 : 0; break;
    case FCPT_PARAM_VISC_TILE: *value = resolve_visc_tile(P); break;
    case FCPT_PARAM_KICK_TILE: *value = resolve_kick_tile(P); break;
    case FCPT_PARAM_SPLIT_ROUNDING: *value = P.split_rounding; break;
    case FCPT_PARAM_REBASE_INTERVAL: *value = ctx->rebase_interval; break;
    default: return fail(ctx, FCPT_ERR_ARG, "unknown launch parameter");
    }
    return FCPT_OK;
}

int fcpt_autotune(fcpt_ctx *ctx, int32_t reps, int32_t *radial_tile, int32_t *radial_waves)
{
    CTX_SYNCED(ctx);
    if (reps <= 0) return fail(ctx, FCPT_ERR_ARG, "reps must be positive");
    int rc = ensure_events(ctx);
    if (rc) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    KParams &P = ctx->P;
    const int tiles[] = {1, 2, 4, 8, 16, 32};
    const int waves[] = {3, 2};
    double best = 1e300;
    int best_t = radial_tile_for(P.nr, P.ns, 0), best_w = 3;
    for (int w : waves) {
        for (int t : tiles) {
            if (t > P.nr) continue;
            P.radial_tile = t;
            P.radial_waves = w;
            launch_radial_sweep(P, ctx->stream);  // warm-up
            HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
            for (int k = 0; k < reps; ++k) launch_radial_sweep(P, ctx->stream);
            HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
            LAUNCH_CHECK(ctx);
            HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
            float ms = 0.f;
            HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            if (ms < best) {
                best = ms;
                best_t = t;
                best_w = w;
            }
        }
    }
    P.radial_tile = best_t;
    P.radial_waves = best_w;
    if (ctx->step_graph) {
        HIP_TRY(ctx, hipGraphExecDestroy(ctx->step_graph));
        ctx->step_graph = nullptr;
    }
    if (radial_tile) *radial_tile = best_t;
    if (radial_waves) *radial_waves = best_w;
    return FCPT_OK;
}

int fcpt_disk_on_planet_accel(fcpt_ctx *ctx, const fcpt_force_body *bodies, int32_t n_bodies,
                              int32_t flags, double *accel_xy)
{
    CTX_SYNCED(ctx);
    if (n_bodies < 0 || n_bodies > kForceMaxBodies || (n_bodies > 0 && (!bodies || !accel_xy)))
        return fail(ctx, FCPT_ERR_ARG, "1 .. %d bodies", kForceMaxBodies);
    if (flags & ~(FCPT_FORCE_NO_STAR_SMOOTHING | FCPT_FORCE_SMOOTHING_PLANETLOC |
                  FCPT_FORCE_CORRECT_SELFGRAVITY))
        return fail(ctx, FCPT_ERR_ARG, "unknown disk-force flags");
    if (n_bodies == 0) return FCPT_OK;
    ForceBodies B{};
    B.n = n_bodies;
    B.flags = flags;
    for (int k = 0; k < n_bodies; ++k) {
        B.x[k] = bodies[k].x;
        B.y[k] = bodies[k].y;
        B.r[k] = bodies[k].r;
        B.r_cubic[k] = bodies[k].r_cubic;
        // compute_smoothing (Force.cpp:145-159): fixed per body, or thickness x H(cell)
        if ((flags & FCPT_FORCE_NO_STAR_SMOOTHING) && k == 0)
            B.smooth_fixed[k] = 0.0;
        else if (flags & FCPT_FORCE_SMOOTHING_PLANETLOC)
            B.smooth_fixed[k] = ctx->desc.thickness_smoothing *
                                (ctx->desc.aspect_ratio * std::pow(bodies[k].r, 1 + ctx->desc.flaring_index));
        else
            B.smooth_fixed[k] = -1.0;
    }
    launch_disk_force(ctx->P, ctx->stream, B, ctx->force_sig1d, ctx->force_part, ctx->force_out);
    LAUNCH_CHECK(ctx);
    HIP_TRY(ctx, hipMemcpyAsync(ctx->force_host, ctx->force_out, 4 * kForceMaxBodies * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < n_bodies; ++k) {
        const double *o = ctx->force_host + 4 * k;
        accel_xy[2 * k] = o[0] + o[2];      // inner + outer, Force.cpp:118-119
        accel_xy[2 * k + 1] = o[1] + o[3];
    }
    return FCPT_OK;
}

// ---- snapshots -------------------------------------------------------------

// the reference's grid names (data.cpp:13-250) of the fields this library holds
static const char *const kGridName[FCPT_NUM_FIELDS] = {
    "Sigma", "vrad", "vazi", "energy", "pressure", "soundspeed", "scale_height", "viscosity",
    "potential", "q_r", "q_phi", "div_v", "tau_r_r", "tau_r_phi", "tau_phi_phi", "Qplus",
    "Qminus", "Temperature", "dens0", "vrad0", "vazi0", "energy0"};

static std::string grid_path(const char *dir, int field)
{
    return std::string(dir) + "/" + kGridName[field] + ".dat";
}

static int write_file(fcpt_ctx *ctx, const std::string &path, const void *data, size_t bytes)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(ctx, FCPT_ERR_ARG, "cannot write %s", path.c_str());
    const size_t n = bytes ? std::fwrite(data, 1, bytes, f) : 0;
    const int closed = std::fclose(f);
    if (n != bytes || closed != 0) return fail(ctx, FCPT_ERR_ARG, "short write to %s", path.c_str());
    return FCPT_OK;
}

// a whole file of exactly `bytes` bytes; 1 if it does not exist
static int read_file(fcpt_ctx *ctx, const std::string &path, void *data, size_t bytes)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return 1;
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (size < 0 || (size_t)size != bytes) {
        std::fclose(f);
        return fail(ctx, FCPT_ERR_ARG, "%s holds %ld bytes, expected %zu", path.c_str(), size, bytes);
    }
    const size_t n = std::fread(data, 1, bytes, f);
    std::fclose(f);
    if (n != bytes) return fail(ctx, FCPT_ERR_ARG, "short read from %s", path.c_str());
    return FCPT_OK;
}

static size_t grid_doubles(const fcpt_ctx *ctx, int field)
{
    return field_rows(field, ctx->P.nr) * (size_t)ctx->P.ns;
}

int fcpt_write_snapshot(fcpt_ctx *ctx, const char *dir, const fcpt_snapshot_info *info, uint32_t fields)
{
    CTX_SYNCED(ctx);
    if (!dir || !info) return fail(ctx, FCPT_ERR_ARG, "null snapshot directory or info");
    if (fields == 0) {
        fields = (1u << FCPT_SIGMA) | (1u << FCPT_VRAD) | (1u << FCPT_VAZI) | (1u << FCPT_ENERGY);
        if (ctx->P.adiabatic) fields |= (1u << FCPT_QPLUS) | (1u << FCPT_QMINUS);
    }
    if (fields >> FCPT_NUM_FIELDS) return fail(ctx, FCPT_ERR_ARG, "unknown field in the mask");
    std::vector<double> host;
    for (int f = 0; f < FCPT_NUM_FIELDS; ++f) {
        if (!(fields & (1u << f))) continue;
        host.resize(grid_doubles(ctx, f));
        int rc = fcpt_download_field(ctx, f, host.data());
        if (!rc) rc = write_file(ctx, grid_path(dir, f), host.data(), host.size() * sizeof(double));
        if (rc) return rc;
    }
    fcpt_snapshot_info m = *info;
    double t = 0.0, last_dt = 0.0;
    int32_t n_monitor = 0;
    int rc = fcpt_get_time(ctx, &t, &last_dt, &n_monitor);
    if (rc) return rc;
    m.time = t;
    m.last_dt = last_dt;
    m.n_monitor = (uint32_t)n_monitor;
    static_assert(sizeof(fcpt_snapshot_info) == 48, "misc.bin record");
    return write_file(ctx, std::string(dir) + "/misc.bin", &m, sizeof(m));
}

int fcpt_read_snapshot(fcpt_ctx *ctx, const char *dir, fcpt_snapshot_info *info)
{
    CTX_SYNCED(ctx);
    if (!dir) return fail(ctx, FCPT_ERR_ARG, "null snapshot directory");
    fcpt_snapshot_info m{};
    int rc = read_file(ctx, std::string(dir) + "/misc.bin", &m, sizeof(m));
    if (rc == 1) return fail(ctx, FCPT_ERR_ARG, "no misc.bin in %s", dir);
    if (rc) return rc;
    const bool adi = ctx->P.adiabatic != 0;
    std::vector<double> g[4];
    const int want[4] = {FCPT_SIGMA, FCPT_VRAD, FCPT_VAZI, FCPT_ENERGY};
    for (int k = 0; k < (adi ? 4 : 3); ++k) {
        g[k].resize(grid_doubles(ctx, want[k]));
        rc = read_file(ctx, grid_path(dir, want[k]), g[k].data(), g[k].size() * sizeof(double));
        if (rc == 1) return fail(ctx, FCPT_ERR_ARG, "no %s", grid_path(dir, want[k]).c_str());
        if (rc) return rc;
    }
    rc = fcpt_upload_state(ctx, g[0].data(), g[1].data(), g[2].data(), adi ? g[3].data() : nullptr);
    if (rc) return rc;
    if (adi) {
        // Qplus / Qminus make the restart bit for bit (restart.cpp:70-86);
        // without them the upload's heating/cooling estimate stays
        for (int f : {FCPT_QPLUS, FCPT_QMINUS}) {
            std::vector<double> q(grid_doubles(ctx, f));
            rc = read_file(ctx, grid_path(dir, f), q.data(), q.size() * sizeof(double));
            if (rc == 0) rc = fcpt_upload_field(ctx, f, q.data());
            else if (rc == 1) rc = FCPT_OK;
            if (rc) return rc;
        }
    }
    rc = fcpt_set_time(ctx, m.time, m.last_dt, (int32_t)m.n_monitor);
    if (rc) return rc;
    if (info) *info = m;
    return fcpt_synchronize(ctx);
}

int fcpt_read_snapshot_reference(fcpt_ctx *ctx, const char *dir)
{
    CTX_CHECK(ctx);
    if (!dir) return fail(ctx, FCPT_ERR_ARG, "null snapshot directory");
    const bool adi = ctx->P.adiabatic != 0;
    const int src[4] = {FCPT_SIGMA, FCPT_VRAD, FCPT_VAZI, FCPT_ENERGY};
    const int dst[4] = {FCPT_SIGMA0, FCPT_VRAD0, FCPT_VAZI0, FCPT_ENERGY0};
    for (int k = 0; k < (adi ? 4 : 3); ++k) {
        std::vector<double> g(grid_doubles(ctx, src[k]));
        int rc = read_file(ctx, grid_path(dir, src[k]), g.data(), g.size() * sizeof(double));
        if (rc == 1) return fail(ctx, FCPT_ERR_ARG, "no %s", grid_path(dir, src[k]).c_str());
        if (!rc) rc = fcpt_upload_field(ctx, dst[k], g.data());
        if (rc) return rc;
    }
    return FCPT_OK;
}

int fcpt_write_grid_files(fcpt_ctx *ctx, const char *outdir, const char *radial_spacing)
{
    CTX_CHECK(ctx);
    if (!outdir || !radial_spacing) return fail(ctx, FCPT_ERR_ARG, "null output directory or spacing");
    std::string text;
    char line[512];
    for (int i = 0; i <= ctx->P.nr; ++i) {
        std::snprintf(line, sizeof(line), "%.18g\n", ctx->Rinf[i]);
        text += line;
    }
    int rc = write_file(ctx, std::string(outdir) + "/used_rad.dat", text.data(), text.size());
    if (rc) return rc;
    text = "#RMIN\tRMAX\tPHIMIN\tPHIMAX          \tNRAD\tNAZ\tNGHRAD\tNGHAZ\tRadial_spacing\n";
    std::snprintf(line, sizeof(line), "%.16g\t%.16g\t%.16g\t%.16g\t%d\t%d\t%d\t%d\t%s\n", ctx->P.rmin,
                  ctx->P.rmax, 0.0, 2 * M_PI, ctx->P.nr, ctx->P.ns, 1, 1, radial_spacing);
    text += line;
    return write_file(ctx, std::string(outdir) + "/dimensions.dat", text.data(), text.size());
}

int fcpt_totals(fcpt_ctx *ctx, double radius, double *mass, double *angular_momentum)
{
    CTX_CHECK(ctx);
    launch_totals(ctx->P, ctx->stream, radius, ctx->totals_ring, ctx->totals_out);
    LAUNCH_CHECK(ctx);
    HIP_TRY(ctx, hipMemcpyAsync(ctx->totals_host, ctx->totals_out, 2 * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (mass) *mass = ctx->totals_host[0];
    if (angular_momentum) *angular_momentum = ctx->totals_host[1];
    return FCPT_OK;
}

int fcpt_get_ring_geometry(fcpt_ctx *ctx, double *rmed, double *surf)
{
    if (!ctx) return FCPT_ERR_ARG;
    const size_t n = (size_t)ctx->P.nr;
    if (rmed) memcpy(rmed, ctx->Rmed.data(), n * sizeof(double));
    if (surf) memcpy(surf, ctx->Surf.data(), n * sizeof(double));
    return FCPT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Totals that allocate nothing.  The ring geometry, the per-ring partial sums
// and the two results live in `scratch`, a file-scope device array: every GPU
// the library runs on holds its own copy until the process ends, at its own
// address, which is looked up once per device with hipGetSymbolAddress; host
// code uses only that address (the array's name is a host-side stand-in).
// Layout for nr rings: Rmed [0, nr), Surf [nr, 2 nr), ring partials
// [2 nr, 6 nr) (four per ring), results [6 nr, 6 nr + 2).  Every call writes
// each slot it later reads (the ring kernel writes all 4 nr partials, the
// fold both results), so nothing has to start at zero.  Calls on one GPU
// share the array: calls that overlap on one GPU race.
constexpr int kNoAllocMaxRings = FCPT_NOALLOC_MAX_RINGS;
static __device__ __align__(256) double scratch[6 * kNoAllocMaxRings + 2];

namespace {
constexpr int kNoAllocMaxDevices = 64;

std::mutex g_scratch_mu;
double *g_scratch_ptr[kNoAllocMaxDevices];
thread_local std::string g_noalloc_err;

int noalloc_fail(int code, const char *fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_noalloc_err = buf;
    return code;
}

// the current device's copy of `scratch`
int scratch_address(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return noalloc_fail(FCPT_ERR_DEVICE, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return noalloc_fail(FCPT_ERR_DEVICE, "device %d beyond the %d devices served", dev, kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return noalloc_fail(FCPT_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return FCPT_OK;
}
}  // namespace

extern "C" {

const char *fcpt_noalloc_last_error(void) { return g_noalloc_err.c_str(); }

int fcpt_totals_noalloc(int32_t nr, int32_t nphi, const double *rmed, const double *surf,
                        const double *sigma, const double *vazi, double omega_frame, double radius,
                        void *stream, double *mass, double *angular_momentum)
{
    g_noalloc_err.clear();
    if (!rmed || !surf || !sigma || !vazi || nr < 3 || nphi < 1)
        return noalloc_fail(FCPT_ERR_ARG, "null pointer or nr < 3 / nphi < 1");
    if (nr > kNoAllocMaxRings)
        return noalloc_fail(FCPT_ERR_ARG, "nr = %d exceeds the %d rings of the static scratch", nr,
                            kNoAllocMaxRings);
    double *s = nullptr;
    int rc = scratch_address(&s);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    double *g_rmed = s, *g_surf = s + nr, *ring = s + 2 * (size_t)nr, *out = s + 6 * (size_t)nr;
    hipError_t e = hipMemcpyAsync(g_rmed, rmed, (size_t)nr * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g_surf, surf, (size_t)nr * sizeof(double), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return noalloc_fail(FCPT_ERR_DEVICE, "geometry upload: %s", hipGetErrorString(e));
    // only the fields k_totals_rings reads: sizes, Rmed, Surf, the frame, Sigma, v_phi
    KParams P{};
    P.nr = nr;
    P.ns = nphi;
    P.Rmed = g_rmed;
    P.Surf = g_surf;
    P.omega_frame = omega_frame;
    P.split = 0;
    P.f[FCPT_SIGMA] = const_cast<double *>(sigma);
    P.f[FCPT_VAZI] = const_cast<double *>(vazi);
    launch_totals(P, st, radius, ring, out);
    e = hipGetLastError();
    if (e != hipSuccess) return noalloc_fail(FCPT_ERR_DEVICE, "kernel launch: %s", hipGetErrorString(e));
    double res[2];
    e = hipMemcpyAsync(res, out, 2 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return noalloc_fail(FCPT_ERR_DEVICE, "result read-back: %s", hipGetErrorString(e));
    if (mass) *mass = res[0];
    if (angular_momentum) *angular_momentum = res[1];
    return FCPT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches: many disks advanced by one launch per stage (plan_batch_step)

struct fcpt_batch {
    int device = 0;
    int split = 0;
    int radial_waves = 3;
    std::vector<fcpt_ctx *> disks;
    std::vector<bool> lazy;  // the disk's batched step forms c_s, nu and T where needed
    KParams *Ps = nullptr;  // device array: every disk's parameters, then their swapped views
    std::vector<BatchStageDev> stages;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t done = nullptr;  // end of the last batched run; every disk's stream waits for it
    std::string err;
    // the parameters the batched kernels see, and which members' radial tile
    // the batch chooses (left to the size rule by the disk)
    std::vector<KParams> host;
    std::vector<bool> auto_radial;
    std::vector<void *> plan_allocs;  // Ps and the stage tables of the current plan
};

namespace {
int bfail(fcpt_batch *b, int code, const char *msg)
{
    if (b) b->err = msg;
    return code;
}
#define BHIP(b, expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            if (b) b->err = std::string(#expr) + ": " + hipGetErrorString(e_);                 \
            return FCPT_ERR_DEVICE;                                                            \
        }                                                                                      \
    } while (0)

void enqueue_batch_step(const fcpt_batch *b, hipStream_t st)
{
    for (const BatchStageDev &s : b->stages) launch_batch_stage(b->Ps, s, b->split, b->radial_waves, st);
}
}  // namespace

extern "C" {

const char *fcpt_batch_last_error(const fcpt_batch *b) { return b ? b->err.c_str() : "null batch"; }

// the device side of the batch's plan from b->host: every disk's parameters
// (disk d at d, its swapped view for the fused passes at n + d) and the work
// table of every stage; replaces the previous plan's arrays
static int upload_batch_plan(fcpt_batch *b)
{
    for (void *p : b->plan_allocs) hipFree(p);
    b->plan_allocs.clear();
    b->stages.clear();
    b->Ps = nullptr;
    auto alloc = [&](size_t bytes) -> void * {
        void *q = nullptr;
        if (hipMalloc(&q, bytes ? bytes : 16) != hipSuccess) return nullptr;
        b->plan_allocs.push_back(q);
        return q;
    };
    std::vector<KParams> views(b->host);
    for (const KParams &P : b->host) views.push_back(swapped_view(P));
    b->Ps = static_cast<KParams *>(alloc(views.size() * sizeof(KParams)));
    if (!b->Ps || hipMemcpy(b->Ps, views.data(), views.size() * sizeof(KParams), hipMemcpyHostToDevice) != hipSuccess)
        return bfail(b, FCPT_ERR_DEVICE, "batch parameters upload");
    for (const BatchStageHost &h : plan_batch_step(b->host)) {
        void *m = alloc(h.entries.size() * sizeof(BatchEntry));
        if (!m || hipMemcpy(m, h.entries.data(), h.entries.size() * sizeof(BatchEntry), hipMemcpyHostToDevice) !=
                      hipSuccess)
            return bfail(b, FCPT_ERR_DEVICE, "batch work table upload");
        b->stages.push_back(
            BatchStageDev{h.kind, h.arg, static_cast<const BatchEntry *>(m), (unsigned)h.entries.size()});
    }
    if (b->graph) {
        hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return FCPT_OK;
}

int fcpt_batch_create(fcpt_ctx *const *disks, int32_t n_disks, fcpt_batch **out)
{
    if (!disks || !out || n_disks <= 0) return FCPT_ERR_ARG;
    *out = nullptr;
    fcpt_batch *b = new fcpt_batch();
    b->device = disks[0] ? disks[0]->device : 0;
    b->split = disks[0] ? disks[0]->P.split : 0;
    std::vector<KParams> host;
    for (int k = 0; k < n_disks; ++k) {
        fcpt_ctx *c = disks[k];
        if (!c || c->device != b->device || c->P.split != b->split ||
            c->P.integrator != FCPT_INTEGRATOR_EULER) {
            delete b;
            return FCPT_ERR_ARG;  // one device, one storage mode, Euler disks
        }
        b->disks.push_back(c);
        // the fields the batched kernels read, stored
        if (c->derived_stale) {
            if (hipSetDevice(c->device) != hipSuccess || sync_derived(c) ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                delete b;
                return FCPT_ERR_DEVICE;
            }
        }
        KParams P = c->P;
        // the radial tile is chosen below, over the whole batch; in split
        // storage the walks' tiles as the disk's own launches choose them
        // (whether the walks run decides where split storage rounds)
        if (P.split) {
            P.kick_tile = resolve_kick_tile(P);
            P.visc_tile = resolve_visc_tile(P);
        }
        host.push_back(P);
    }
    // FP64 storage: the walks' tiles of the members that leave them to the
    // size rule, chosen over the batch's wavefronts as the radial tile below
    // (a member too small to fill the device alone still runs the walks;
    // FP64 results do not depend on the tiles or on whether the walks run)
    if (!b->split) {
        auto waves = [&](int t) {
            long w = 0;
            for (const KParams &P : host) w += walk_waves(P, P.visc_tile > 0 ? P.visc_tile : t);
            return w;
        };
        int t = 8;
        while (t > 2 && waves(t) < 3072) t /= 2;
        for (KParams &P : host) {
            if (P.kick_tile <= 0) P.kick_tile = t;
            if (P.visc_tile <= 0) P.visc_tile = t;
        }
    }
    for (KParams &P : host) {
        // the batch runs the fused step: c_s, nu and T formed where needed
        // as in the disk's own fused step
        const bool walks = P.adiabatic && kick_walk_applies(P);
        P.e_partner = walks && P.split ? 1 : 0;
        P.lazy_derived = walks ? 1 : 0;
        b->lazy.push_back(walks);
    }
    // The radial sweep's tile of the members that leave it to the size rule.
    // A single disk's rule shrinks the tile until that disk alone fills the
    // device with wavefronts; in a batch all members' blocks share one
    // launch, so the tile is the largest that fills the device with the
    // batch's wavefronts (fewer tiles, hence fewer window reloads at tile
    // starts; the sweep's results do not depend on the tile).
    for (const KParams &P : host) b->auto_radial.push_back(P.radial_tile <= 0);
    {
        auto waves = [&](int t) {
            long w = 0;
            for (const KParams &P : host)
                w += (long)((P.ns + 63) / 64) * ((P.nr + (P.radial_tile > 0 ? P.radial_tile : t) - 1) /
                                                  (P.radial_tile > 0 ? P.radial_tile : t));
            return w;
        };
        int t = 32;
        while (t > 2 && waves(t) < 3072) t /= 2;
        for (KParams &P : host)
            if (P.radial_tile <= 0) P.radial_tile = t;
    }
    b->host = host;
    int rc = hipSetDevice(b->device) == hipSuccess ? FCPT_OK : FCPT_ERR_DEVICE;
    if (!rc) rc = upload_batch_plan(b);
    if (!rc && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) rc = FCPT_ERR_DEVICE;
    if (!rc && hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) rc = FCPT_ERR_DEVICE;
    if (rc) {
        fcpt_batch_destroy(b);
        return rc;
    }
    *out = b;
    return FCPT_OK;
}

int fcpt_batch_destroy(fcpt_batch *b)
{
    if (!b) return FCPT_OK;
    hipSetDevice(b->device);
    if (b->stream) {
        hipStreamSynchronize(b->stream);
        hipStreamDestroy(b->stream);
    }
    if (b->graph) hipGraphExecDestroy(b->graph);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->done) hipEventDestroy(b->done);
    for (void *p : b->plan_allocs) hipFree(p);
    delete b;
    return FCPT_OK;
}

int fcpt_batch_set_radial_waves(fcpt_batch *b, int32_t waves)
{
    if (!b) return FCPT_ERR_ARG;
    if (waves != 2 && waves != 3) return bfail(b, FCPT_ERR_ARG, "radial waves must be 2 or 3");
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipStreamSynchronize(b->stream));
    b->radial_waves = waves;
    if (b->graph) {
        BHIP(b, hipGraphExecDestroy(b->graph));
        b->graph = nullptr;
    }
    return FCPT_OK;
}

int fcpt_batch_synchronize(fcpt_batch *b)
{
    if (!b) return FCPT_ERR_ARG;
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipStreamSynchronize(b->stream));
    return FCPT_OK;
}

int fcpt_batch_run_steps(fcpt_batch *b, int32_t n_steps, int32_t use_graph)
{
    if (!b) return FCPT_ERR_ARG;
    if (n_steps < 0) return bfail(b, FCPT_ERR_ARG, "negative step count");
    BHIP(b, hipSetDevice(b->device));
    // the disks' own queued work (uploads, start-up calls) comes first
    for (fcpt_ctx *c : b->disks) BHIP(b, hipStreamSynchronize(c->stream));
    if (use_graph && !b->graph) {
        hipGraph_t g = nullptr;
        BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        enqueue_batch_step(b, b->stream);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (e != hipSuccess) return bfail(b, FCPT_ERR_DEVICE, hipGetErrorString(e));
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) {
            b->graph = nullptr;
            return bfail(b, FCPT_ERR_DEVICE, hipGetErrorString(e));
        }
    }
    for (int left = n_steps; left > 0;) {
        // up to the next background move of a member in split storage
        int m = left;
        for (fcpt_ctx *c : b->disks)
            if (periodic_rebase(c)) m = std::min(m, c->rebase_interval - c->steps_since_rebase);
        if (use_graph) {
            for (int k = 0; k < m; ++k) BHIP(b, hipGraphLaunch(b->graph, b->stream));
        } else {
            for (int k = 0; k < m; ++k) enqueue_batch_step(b, b->stream);
            BHIP(b, hipGetLastError());
        }
        left -= m;
        // a disk call made after this one (a download, the totals) is ordered
        // after the batched steps on the disk's own stream
        BHIP(b, hipEventRecord(b->done, b->stream));
        for (fcpt_ctx *c : b->disks) BHIP(b, hipStreamWaitEvent(c->stream, b->done, 0));
        for (size_t k = 0; k < b->disks.size(); ++k)
            if (b->lazy[k]) b->disks[k]->derived_stale = true;
        for (fcpt_ctx *c : b->disks) {
            if (!periodic_rebase(c) || (c->steps_since_rebase += m) < c->rebase_interval) continue;
            if (rebase_backgrounds(c)) return bfail(b, FCPT_ERR_DEVICE, c->err.c_str());
        }
    }
    return FCPT_OK;
}

int fcpt_batch_time_kernel(fcpt_batch *b, int32_t kernel, int32_t reps, double *ms_per_launch)
{
    if (!b) return FCPT_ERR_ARG;
    if (reps <= 0) return bfail(b, FCPT_ERR_ARG, "reps must be positive");
    int kinds[2] = {-1, -1};
    switch (kernel) {
    case FCPT_KERNEL_RADIAL_SWEEP: kinds[0] = B_RADIAL_ISO; kinds[1] = B_RADIAL_ADI; break;
    case FCPT_KERNEL_AZIMUTHAL: kinds[0] = B_AZIM_ISO; kinds[1] = B_AZIM_ADI; break;
    case FCPT_KERNEL_CFL: kinds[0] = B_CFL_RINGS; break;
    case FCPT_KERNEL_POTENTIAL: kinds[0] = B_POTENTIAL; break;
    case FCPT_KERNEL_STRESS:
    case FCPT_KERNEL_VISC_FUSED:
        // the batched step runs the fused viscous pass, on views that differ
        // between disks; timed alone it could write the fields
        return bfail(b, FCPT_ERR_ARG, "the viscous kernels are not timed on a batch");
    default: return bfail(b, FCPT_ERR_ARG, "unknown kernel");
    }
    BHIP(b, hipSetDevice(b->device));
    if (!b->ev0) BHIP(b, hipEventCreate(&b->ev0));
    if (!b->ev1) BHIP(b, hipEventCreate(&b->ev1));
    auto one = [&]() {
        for (const BatchStageDev &s : b->stages) {
            if (s.kind != kinds[0] && s.kind != kinds[1]) continue;
            launch_batch_stage(b->Ps, s, b->split, b->radial_waves, b->stream);
        }
    };
    BHIP(b, hipStreamSynchronize(b->stream));
    one();
    BHIP(b, hipEventRecord(b->ev0, b->stream));
    for (int k = 0; k < reps; ++k) one();
    BHIP(b, hipEventRecord(b->ev1, b->stream));
    BHIP(b, hipGetLastError());
    BHIP(b, hipEventSynchronize(b->ev1));
    float ms = 0.f;
    BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (ms_per_launch) *ms_per_launch = (double)ms / reps;
    return FCPT_OK;
}

// The batch's radial tile (for the members that leave it to the size rule)
// and register budget by the radial stage's time: each candidate's plan is
// uploaded and its radial stage timed over `reps` launches, which write the
// transport's scratch arrays only, so the disks' state is untouched.
int fcpt_batch_autotune(fcpt_batch *b, int32_t reps, int32_t *radial_tile, int32_t *radial_waves)
{
    if (!b) return FCPT_ERR_ARG;
    if (reps <= 0) return bfail(b, FCPT_ERR_ARG, "reps must be positive");
    BHIP(b, hipSetDevice(b->device));
    BHIP(b, hipStreamSynchronize(b->stream));
    int cur_t = 0;
    for (size_t k = 0; k < b->host.size(); ++k)
        if (b->auto_radial[k]) cur_t = b->host[k].radial_tile;
    int best_t = cur_t, best_w = b->radial_waves;
    double best = 1e300;
    const int tiles[] = {4, 8, 16, 32};
    for (int w : {3, 2}) {
        for (int t : tiles) {
            if (cur_t == 0 && t != tiles[0]) break;  // no member's tile is the batch's to choose
            for (size_t k = 0; k < b->host.size(); ++k)
                if (b->auto_radial[k]) b->host[k].radial_tile = t;
            int rc = upload_batch_plan(b);
            b->radial_waves = w;
            double ms = 0.0;
            if (!rc) rc = fcpt_batch_time_kernel(b, FCPT_KERNEL_RADIAL_SWEEP, reps, &ms);
            if (rc) return rc;
            if (ms < best) {
                best = ms;
                best_t = t;
                best_w = w;
            }
        }
    }
    for (size_t k = 0; k < b->host.size(); ++k)
        if (b->auto_radial[k]) b->host[k].radial_tile = best_t;
    b->radial_waves = best_w;
    const int rc = upload_batch_plan(b);
    if (rc) return rc;
    if (radial_tile) *radial_tile = best_t;
    if (radial_waves) *radial_waves = best_w;
    return FCPT_OK;
}

int fcpt_batch_time_steps(fcpt_batch *b, int32_t n_steps, int32_t use_graph, double *ms_total)
{
    if (!b) return FCPT_ERR_ARG;
    BHIP(b, hipSetDevice(b->device));
    if (!b->ev0) BHIP(b, hipEventCreate(&b->ev0));
    if (!b->ev1) BHIP(b, hipEventCreate(&b->ev1));
    int rc = fcpt_batch_run_steps(b, 0, use_graph);  // capture outside the timed region
    if (rc) return rc;
    BHIP(b, hipStreamSynchronize(b->stream));
    BHIP(b, hipEventRecord(b->ev0, b->stream));
    rc = fcpt_batch_run_steps(b, n_steps, use_graph);
    if (rc) return rc;
    BHIP(b, hipEventRecord(b->ev1, b->stream));
    BHIP(b, hipEventSynchronize(b->ev1));
    float ms = 0.f;
    BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    for (fcpt_ctx *c : b->disks) {
        DevScalars sc;
        BHIP(b, hipMemcpy(&sc, c->s_dev, sizeof(DevScalars), hipMemcpyDeviceToHost));
        if (sc.error_flags & 1) {
            // reported once, like a disk's own calls report it
            BHIP(b, hipMemset(&c->s_dev->error_flags, 0, sizeof(int)));
            return bfail(b, FCPT_ERR_NONFINITE, "a disk's CFL time step became non-finite");
        }
    }
    if (ms_total) *ms_total = (double)ms;
    return FCPT_OK;
}

}  // extern "C"
