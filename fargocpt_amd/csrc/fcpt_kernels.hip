// gfx950 kernels of the FargoCPT hydro step.
//
// Every kernel restates one loop nest of the reference (file:line given above
// it) with the same expression order; the library is compiled with
// -ffp-contract=off so that add/mul/div/sqrt round exactly like the
// reference's host build.  Threads map x -> azimuthal index j (contiguous in
// HBM, so a wavefront reads 64 consecutive values) and y -> ring index i.
//
// Storage: every kernel is instantiated twice.  SPLIT = false keeps each field
// in FP64.  SPLIT = true keeps each 2-D field as an FP32 deviation from an FP64
// per-ring background (value = bg[ring] + (double)deviation), which halves the
// HBM traffic; the arithmetic is FP64 in both instantiations and the FP64
// instantiation performs exactly the operations of the reference.
#include <type_traits>

#include "fcpt_device.hpp"
#include "fcpt_kernels.hpp"

namespace fcpt {

#define IX(i, j) ((size_t)(i) * (size_t)ns + (size_t)(j))

__device__ __forceinline__ double std_min(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double std_max(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double van_leer_lim(double a, double b)
{
    // TransportEuler.cpp:306-312
    if (a * b > 0.0) return 2.0 * a * b / (a + b);
    return 0;
}
__device__ __forceinline__ double omega_kepler(const KParams &P, double r)
{
    return sqrt(P.G * P.mstar / (r * r * r));  // Theo.cpp:246-249
}

// field accessor: an FP64 array, or FP32 deviations from an FP64 ring background
template <bool SPLIT>
struct Fld {
    using T = typename std::conditional<SPLIT, float, double>::type;
    T *p;
    const double *bg;
    int ns;
    __device__ __forceinline__ double operator()(int i, int j) const
    {
        const size_t c = (size_t)i * (size_t)ns + (size_t)j;
        if constexpr (SPLIT) return bg[i] + (double)p[c];
        else return p[c];
    }
    __device__ __forceinline__ void set(int i, int j, double v) const
    {
        const size_t c = (size_t)i * (size_t)ns + (size_t)j;
        if constexpr (SPLIT) p[c] = (float)(v - bg[i]);
        else p[c] = v;
    }
};

template <bool SPLIT>
__device__ __forceinline__ Fld<SPLIT> fld(const KParams &P, int f)
{
    return Fld<SPLIT>{static_cast<typename Fld<SPLIT>::T *>(P.f[f]), P.bg[f], P.ns};
}
template <bool SPLIT>
__device__ __forceinline__ Fld<SPLIT> buf(const KParams &P, void *p, const double *bg)
{
    return Fld<SPLIT>{static_cast<typename Fld<SPLIT>::T *>(p), bg, P.ns};
}

#define FLD(name, id) const Fld<SPLIT> name = fld<SPLIT>(P, id)

// 2-D launch over rows [r0, r1) x all sectors
#define CELL_LOOP_HEAD(r0, r1)                                   \
    const int ns = P.ns;                                         \
    const int j = blockIdx.x * blockDim.x + threadIdx.x;         \
    const int i = (r0) + blockIdx.y * blockDim.y + threadIdx.y;  \
    if (j >= ns || i >= (r1)) return;

// ---------------------------------------------------------------------------
// ring sums (cfl.cpp:196-205, TransportEuler.cpp:174-190): one block per ring,
// pairwise tree order
template <int BS>
__device__ double block_sum(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = sh[t] + sh[t + s];
        __syncthreads();
    }
    return sh[0];
}

// mode 0: CFL ring mean
// mode 1: transport ring mean + FARGO shift + uniform residual
//         (TransportEuler.cpp:207-236)
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_ring_mean(KParams P, int mode)
{
    __shared__ double sh[256];
    const int ns = P.ns;
    const int i = blockIdx.x;
    FLD(v, FCPT_VAZI);
    double acc = 0.0;
    for (int j = threadIdx.x; j < ns; j += 256) acc += v(i, j);
    const double sum = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) {
        const double vm = sum / (double)ns;
        P.vmean[i] = vm;
        if (mode == 1) {
            const double dt = P.s->dt;
            const double invdt = 1.0 / dt;
            const double Ntilde = vm * P.InvRmed[i] * dt * P.invdphi;
            const double Nround = floor(Ntilde + 0.5);
            P.nshift[i] = (int)Nround;
            P.vuni[i] = (Ntilde - Nround) * P.Rmed[i] * invdt * P.dphi;
        }
    }
}

// cfl.cpp:210-376, per cell dt and the per-ring shear limit.  Each thread
// walks rows with a grid stride; the block minimum (wavefront shuffles, then
// LDS) goes to its own slot of P.ring_min and k_cfl_finalize folds the slots,
// so the result is deterministic and free of atomics.
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_cfl_cells(KParams P)
{
    const int ns = P.ns;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double cand = __longlong_as_double(0x7FF0000000000000ll);
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(cs, FCPT_SOUNDSPEED);
    FLD(nu, FCPT_VISCOSITY);
    const double CFL = P.cfl, dphi = P.dphi;
    const double *vm = P.vmean;
    const double fac2 = P.artvisc_factor * P.artvisc_factor;
    for (int i = 1 + blockIdx.y * blockDim.y + threadIdx.y; i < P.nr - 1 && j < ns;
         i += gridDim.y * blockDim.y) {
        if (j == 0) {
            const double denom =
                fabs(vm[i] * P.InvRmed[i] - vm[i + 1] * P.InvRmed[i + 1]) + 1.0e-100;
            cand = std_min(cand, CFL * dphi / denom);
            if (i == 1) {
                const double denom0 =
                    fabs(vm[0] * P.InvRmed[0] - vm[1] * P.InvRmed[1]) + 1.0e-100;
                cand = std_min(cand, CFL * dphi / denom0);
            }
        }
        const double dxRadial = P.Rsup[i] - P.Rinf[i];
        const double dxAzimuthal = P.Rmed[i] * dphi;
        const double cell_size = std_min(dxRadial, dxAzimuthal);
        const int jn = j == ns - 1 ? 0 : j + 1;
        const double vr0 = vr(i, j), vr1 = vr(i + 1, j), va0 = vaz(i, j), va1 = vaz(i, jn);
        const double v_residual = va0 - vm[i];
        const double invdt1 = cs(i, j) / cell_size;
        const double invdt2 = vr0 / dxRadial;
        const double invdt3 = v_residual / dxAzimuthal;
        double invdt4;
        if (P.artvisc == FCPT_ARTVISC_SN) {
            double dvRadial = vr1 - vr0;
            double dvAzimuthal = va1 - va0;
            if (dvRadial > 0.0) dvRadial = 0.0;
            else dvRadial = -dvRadial;
            if (dvAzimuthal > 0.0) dvAzimuthal = 0.0;
            else dvAzimuthal = -dvAzimuthal;
            invdt4 = 4.0 * fac2 * std_max(dvRadial / dxRadial, dvAzimuthal / dxAzimuthal) * 1.0;
        } else {
            const double eps_rr = (vr1 - vr0) * P.InvDiffRsup[i];
            const double eps_pp = P.InvRmed[i] * ((va1 - va0) * P.invdphi + 0.5 * (vr1 + vr0));
            const double mdiv_V = -std_min(eps_rr + eps_pp, 0.0);
            invdt4 = 4.0 * fac2 * mdiv_V * 1.0;
        }
        const double invdt5 = 4.0 * nu(i, j) / (cell_size * cell_size) * 1.0;
        double invdt6 = 0.0;
        if (P.adiabatic) {
            FLD(qp, FCPT_QPLUS);
            FLD(qm, FCPT_QMINUS);
            FLD(E, FCPT_ENERGY);
            const double inv_limit = 1.0 / P.heating_cooling_cfl_limit;
            invdt6 = inv_limit * fabs((qp(i, j) - qm(i, j)) / E(i, j)) * 1.0;
        }
        const double dt_cell =
            CFL / sqrt(invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 + invdt4 * invdt4 +
                       invdt5 * invdt5 + invdt6 * invdt6);
        // NaN never wins a comparison, exactly like the reference's "if (dt_cell < dt_core)"
        cand = std_min(cand, dt_cell);
    }
    for (int off = 32; off > 0; off >>= 1) cand = std_min(cand, __shfl_xor(cand, off, 64));
    __shared__ double wmin[4];
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    if ((t & 63) == 0) wmin[t >> 6] = cand;
    __syncthreads();
    if (t == 0) {
        const double m = std_min(std_min(wmin[0], wmin[1]), std_min(wmin[2], wmin[3]));
        P.ring_min[blockIdx.y * gridDim.x + blockIdx.x] = m;
    }
}

// mode 0: condition_cfl only; 1: sim::CalculateTimeStep (simulation.cpp:100-118);
// 2: CalculateTimeStep + the monitor clip of sim::run (simulation.cpp:526-542)
__global__ void __launch_bounds__(256) k_cfl_finalize(KParams P, int mode, int nblocks)
{
    __shared__ double sh[256];
    double m = __longlong_as_double(0x7FF0000000000000ll);
    for (int k = threadIdx.x; k < nblocks; k += 256) m = std_min(m, P.ring_min[k]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = std_min(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    DevScalars *s = P.s;
    const double cfl_dt = sh[0];
    s->cfl_dt = cfl_dt;
    if (!(cfl_dt > 0.0) || !isfinite(cfl_dt)) s->error_flags |= 1;
    if (mode >= 1) {
        const double rv = std_min(P.cfl_max_var * s->last_dt, cfl_dt);
        s->last_dt = rv;
        s->cfl_dt = rv;
        if (mode == 2) {
            const double time_next_monitor = (s->n_monitor + 1) * P.monitor_timestep;
            const double time_left_till_write = time_next_monitor - s->time;
            const bool overshoot = rv > time_left_till_write;
            const double dt_stretch_factor = 0.05;
            const bool almost_there = time_left_till_write < rv * (1 + dt_stretch_factor);
            s->dt = (overshoot || almost_there) ? time_left_till_write : rv;
        }
    }
}

__global__ void k_set_dt(KParams P, double dt) { P.s->dt = dt; }
__global__ void k_set_clock(KParams P, double time) { P.s->time = time; }

__global__ void k_set_time(KParams P, double time, double last_dt, int n_monitor)
{
    P.s->time = time;
    P.s->last_dt = last_dt;
    P.s->n_monitor = n_monitor;
}

__global__ void k_set_params(KParams P, fcpt_step_params sp)
{
    DevScalars *s = P.s;
    s->n_planets = sp.n_planets;
    for (int k = 0; k < sp.n_planets; ++k) {
        s->planet[k].x = sp.planet[k].x;
        s->planet[k].y = sp.planet[k].y;
        s->planet[k].mass = sp.planet[k].mass;
        s->planet[k].r_cubic = sp.planet[k].r_cubic;
    }
    s->indirect_x = sp.indirect_x;
    s->indirect_y = sp.indirect_y;
}

// ---------------------------------------------------------------------------
// CalculateNbodyPotential, Pframeforce.cpp:44-85 (+ compute_smoothing, Force.cpp:145-159)
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_potential(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(Hf, FCPT_SCALE_HEIGHT);
    const DevScalars *s = P.s;
    const double x = P.Rmed[i] * P.cosphi[j];
    const double y = P.Rmed[i] * P.sinphi[j];
    const double H = Hf(i, j);
    double acc = 0.0;
    const int np = s->n_planets;
    for (int k = 0; k < np; ++k) {
        const double smooth = P.thickness_smoothing * H;
        const double dx = x - s->planet[k].x;
        const double dy = y - s->planet[k].y;
        const double dist_2 = dx * dx + dy * dy;
        const double d_smoothed = sqrt(dist_2 + smooth * smooth);
        double smooth_factor_klahr = 1.0;
        const double r_sm = s->planet[k].r_cubic;
        if (r_sm > 0.0 && d_smoothed < r_sm) {
            const double q = d_smoothed / r_sm;
            smooth_factor_klahr = (pow(q, 4.0) - 2.0 * pow(q, 3.0) + 2.0 * d_smoothed / r_sm);
        }
        acc += -P.G * s->planet[k].mass / d_smoothed * smooth_factor_klahr;
    }
    acc += -s->indirect_x * x - s->indirect_y * y;
    if constexpr (SPLIT) {
        // The potential is rewritten completely every step, so its background
        // is chosen here: every body seen from distance sqrt(R^2 + |r_k|^2)
        // with the ring's background smoothing.  Each thread evaluates the
        // same expression; the ring's first sector publishes it.
        const double R = P.Rmed[i];
        const double smooth = P.thickness_smoothing * P.bg[FCPT_SCALE_HEIGHT][i];
        double b = 0.0;
        for (int k = 0; k < np; ++k) {
            const double px = s->planet[k].x, py = s->planet[k].y;
            b += -P.G * s->planet[k].mass / sqrt(R * R + (px * px + py * py) + smooth * smooth);
        }
        if (j == 0) P.bg[FCPT_POTENTIAL][i] = b;
        static_cast<float *>(P.f[FCPT_POTENTIAL])[IX(i, j)] = (float)(acc - b);
    } else {
        fld<SPLIT>(P, FCPT_POTENTIAL).set(i, j, acc);
    }
}

// ---------------------------------------------------------------------------
// momentum_update_radial, SourceEuler.cpp:325-372
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_src_vr(KParams P)
{
    CELL_LOOP_HEAD(2, P.nr - 1)
    FLD(S, FCPT_SIGMA);
    FLD(Pr, FCPT_PRESSURE);
    FLD(Phi, FCPT_POTENTIAL);
    FLD(vaz, FCPT_VAZI);
    FLD(vr, FCPT_VRAD);
    const double dt = P.s->dt;
    double gradp = 2.0 / (S(i, j) + S(i - 1, j));
    gradp *= (Pr(i, j) - Pr(i - 1, j));
    gradp *= P.InvDiffRmed[i];
    const double gradphi = (Phi(i, j) - Phi(i - 1, j)) * P.InvDiffRmed[i];
    const int jn = j == ns - 1 ? 0 : j + 1;
    const double vsum = vaz(i, j) + vaz(i, jn) + vaz(i - 1, j) + vaz(i - 1, jn);
    const double vt = 0.25 * vsum + P.Rinf[i] * P.omega_frame;
    const double vt2 = vt * vt;
    const double centrifugal_accel = vt2 * P.InvRinf[i];
    vr.set(i, j, vr(i, j) + dt * (-gradp - gradphi + centrifugal_accel));
}

// momentum_update_azimuthal, SourceEuler.cpp:375-428
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_src_vphi(KParams P)
{
    CELL_LOOP_HEAD(1, P.nr - 1)
    FLD(S, FCPT_SIGMA);
    FLD(Pr, FCPT_PRESSURE);
    FLD(Phi, FCPT_POTENTIAL);
    FLD(vaz, FCPT_VAZI);
    const double dt = P.s->dt;
    const double invdxtheta = 2.0 / (P.dphi * (P.Rsup[i] + P.Rinf[i]));
    const int jp = j == 0 ? ns - 1 : j - 1;
    const double gradp = 2.0 / (S(i, j) + S(i, jp)) * (Pr(i, j) - Pr(i, jp)) * invdxtheta;
    const double gradphi = (Phi(i, j) - Phi(i, jp)) * invdxtheta;
    vaz.set(i, j, vaz(i, j) + dt * (-gradp - gradphi));
}

// compression_heating, SourceEuler.cpp:459-493
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_src_energy(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr - 1)
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(E, FCPT_ENERGY);
    const double dt = P.s->dt;
    const int jn = j == ns - 1 ? 0 : j + 1;
    const double DIV_V = (vr(i + 1, j) * P.Rinf[i + 1] - vr(i, j) * P.Rinf[i]) * P.InvDiffRsupRb[i] +
                         (vaz(i, jn) - vaz(i, j)) * P.invdphi * P.InvRmed[i];
    const double energy_old = E(i, j);
    E.set(i, j, energy_old * exp(-(P.gamma - 1.0) * dt * DIV_V));
}

// assure_temperature_range for one cell, SourceEuler.cpp:136-202
__device__ __forceinline__ double temperature_floor(const KParams &P, double e, double sigma)
{
    const double minimum_energy = P.t_min * sigma / P.mu * P.Rgas / (P.gamma - 1.0);
    const double maximum_energy = P.t_max * sigma / P.mu * P.Rgas / (P.gamma - 1.0);
    if (!(e > minimum_energy)) e = P.t_min * sigma / P.mu * P.Rgas / (P.gamma - 1.0);
    if (!(e < maximum_energy)) e = P.t_max * sigma / P.mu * P.Rgas / (P.gamma - 1.0);
    return e;
}

template <bool SPLIT>
__global__ void __launch_bounds__(256) k_temperature_floor(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(E, FCPT_ENERGY);
    FLD(S, FCPT_SIGMA);
    E.set(i, j, temperature_floor(P, E(i, j), S(i, j)));
}

// ---------------------------------------------------------------------------
// compute_sound_speed / compute_scale_height / update_viscosity for one cell
// (SourceEuler.cpp:957-995, 1121-1154; viscosity.cpp:98-137)
template <bool SPLIT>
__device__ __forceinline__ void derived_cell(const KParams &P, int i, int j, bool with_temperature,
                                            bool with_pressure)
{
    FLD(cs, FCPT_SOUNDSPEED);
    FLD(Hf, FCPT_SCALE_HEIGHT);
    double c_s, h;
    if (P.adiabatic) {
        FLD(E, FCPT_ENERGY);
        FLD(S, FCPT_SIGMA);
        const double e = E(i, j);
        const double sig = S(i, j);
        if (with_temperature) {
            const double c_v_inv = P.mu / P.Rgas * (P.gamma - 1.0);
            fld<SPLIT>(P, FCPT_TEMPERATURE).set(i, j, c_v_inv * e / sig);
        }
        const double g = P.gamma;
        c_s = sqrt(g * (g - 1.0) * e / sig);
        cs.set(i, j, c_s);
        const double inv_omega_kepler = 1.0 / omega_kepler(P, P.Rmed[i]);
        h = c_s / (sqrt(g)) * inv_omega_kepler;
        Hf.set(i, j, h);
        if (with_pressure) fld<SPLIT>(P, FCPT_PRESSURE).set(i, j, (P.gamma - 1.0) * e);
    } else {
        c_s = cs(i, j);
        h = Hf(i, j);
        if (with_pressure) {
            FLD(S, FCPT_SIGMA);
            fld<SPLIT>(P, FCPT_PRESSURE).set(i, j, S(i, j) * (c_s * c_s));
        }
    }
    const double alpha = P.alpha;
    fld<SPLIT>(P, FCPT_VISCOSITY).set(i, j, alpha * h * c_s);
}

// ---------------------------------------------------------------------------
// update_with_artificial_viscosity_TW, artificial_viscosity.cpp:35-88
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_tw_q(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(S, FCPT_SIGMA);
    const double dt = P.s->dt;
    const int jn = j == ns - 1 ? 0 : j + 1;
    const double vr0 = vr(i, j), vr1 = vr(i + 1, j);
    const double eps_rr = (vr1 - vr0) * P.InvDiffRsup[i];
    const double eps_pp = P.InvRmed[i] * ((vaz(i, jn) - vaz(i, j)) * P.invdphi + 0.5 * (vr1 + vr0));
    const double div_V = std_min(eps_rr + eps_pp, 0.0);
    const double Dr = P.Rinf[i + 1] - P.Rinf[i];
    const double rDphi = P.Rmed[i] * P.dphi;
    const double dxm = (ns <= 16) ? std_min(Dr, rDphi) : std_max(Dr, rDphi);
    const double dx_sq = dxm * dxm;
    const double l_sq = (P.artvisc_factor * P.artvisc_factor) * dx_sq;
    const double sig = S(i, j);
    fld<SPLIT>(P, FCPT_Q_R).set(i, j, l_sq * sig * -div_V * (eps_rr - 1.0 / 3.0 * div_V));
    fld<SPLIT>(P, FCPT_Q_PHI).set(i, j, l_sq * sig * -div_V * (eps_pp - 1.0 / 3.0 * div_V));
    if (P.adiabatic && P.artvisc_dissipation) {
        FLD(E, FCPT_ENERGY);
        double e = E(i, j);
        if (i > 1 && i < P.nr - 1) {
            const double d = eps_rr - eps_pp;
            const double Qplus =
                -l_sq * div_V * sig * 1.0 / 3.0 * (eps_rr * eps_rr + eps_pp * eps_pp + d * d);
            e += Qplus * dt;
        }
        // SetTemperatureFloorCeilValues after the art. viscosity (artificial_viscosity.cpp:19-21)
        E.set(i, j, temperature_floor(P, e, sig));
    }
}

// artificial_viscosity.cpp:90-139
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_tw_vel(KParams P, int with_viscosity)
{
    CELL_LOOP_HEAD(0, P.nr)
    // in the fused step, recalculate_viscosity (SourceEuler.cpp:205-223), which is
    // pointwise in Sigma and E and follows this kernel, is done here for every ring
    if (with_viscosity) derived_cell<SPLIT>(P, i, j, false, false);
    if (i < 1 || i >= P.nr - 1) return;
    FLD(S, FCPT_SIGMA);
    FLD(Qr, FCPT_Q_R);
    FLD(Qp, FCPT_Q_PHI);
    FLD(vaz, FCPT_VAZI);
    const double dt = P.s->dt;
    const int jp = j == 0 ? ns - 1 : j - 1;
    const double sigma_phi_avg = 0.5 * (S(i, j) + S(i, jp));
    const double dVp = 2.0 * dt / ((P.Rsup[i] + P.Rinf[i]) * sigma_phi_avg) *
                       (Qp(i, j) - Qp(i, jp)) * P.invdphi;
    vaz.set(i, j, vaz(i, j) + dVp);
    if (i >= 2) {
        FLD(vr, FCPT_VRAD);
        const double sigma_r_avg = 0.5 * (S(i, j) + S(i - 1, j));
        const double r0 = P.Rmed[i], r1 = P.Rmed[i - 1];
        const double dVr = P.radial_viscosity_factor * dt / sigma_r_avg * 2.0 / (r0 * r0 - r1 * r1) *
                           ((Qr(i, j) * r0 - Qr(i - 1, j) * r1) -
                            0.5 * (Qp(i, j) + Qp(i - 1, j)) * (r0 - r1));
        vr.set(i, j, vr(i, j) + dVr);
    }
}

// update_with_artificial_viscosity_SN, artificial_viscosity.cpp:148-211
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_sn_q(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(S, FCPT_SIGMA);
    const double fac2 = P.artvisc_factor * P.artvisc_factor;
    const double sig = S(i, j);
    const double dv_r = vr(i + 1, j) - vr(i, j);
    const double qr = dv_r < 0.0 ? fac2 * sig * (dv_r * dv_r) : 0.0;
    const int jn = j == ns - 1 ? 0 : j + 1;
    const double dv_phi = vaz(i, jn) - vaz(i, j);
    const double qp = dv_phi < 0.0 ? fac2 * sig * (dv_phi * dv_phi) : 0.0;
    fld<SPLIT>(P, FCPT_Q_R).set(i, j, qr);
    fld<SPLIT>(P, FCPT_Q_PHI).set(i, j, qp);
    if (P.adiabatic && P.artvisc_dissipation) {
        FLD(E, FCPT_ENERGY);
        double e = E(i, j);
        if (i >= 1 && i < P.nr - 1) {
            const double dt = P.s->dt;
            const double dxtheta = P.dphi * P.Rmed[i];
            const double invdxtheta = 1.0 / dxtheta;
            e = e - dt * qr * dv_r * P.InvDiffRsup[i] - dt * qp * dv_phi * invdxtheta;
        }
        E.set(i, j, temperature_floor(P, e, sig));
    }
}

// artificial_viscosity.cpp:214-248
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_sn_vel(KParams P, int with_viscosity)
{
    CELL_LOOP_HEAD(0, P.nr)
    // in the fused step, recalculate_viscosity (SourceEuler.cpp:205-223), which is
    // pointwise in Sigma and E and follows this kernel, is done here for every ring
    if (with_viscosity) derived_cell<SPLIT>(P, i, j, false, false);
    if (i < 1 || i >= P.nr - 1) return;
    FLD(S, FCPT_SIGMA);
    FLD(Qr, FCPT_Q_R);
    FLD(Qp, FCPT_Q_PHI);
    const double dt = P.s->dt;
    if (i >= 2) {
        FLD(vr, FCPT_VRAD);
        vr.set(i, j, vr(i, j) - dt * 2.0 / (S(i, j) + S(i - 1, j)) * (Qr(i, j) - Qr(i - 1, j)) *
                                    P.InvDiffRmed[i]);
    }
    const double dxtheta = P.dphi * P.Rmed[i];
    const double invdxtheta = 1.0 / dxtheta;
    const int jp = j == 0 ? ns - 1 : j - 1;
    FLD(vaz, FCPT_VAZI);
    vaz.set(i, j, vaz(i, j) - dt * 2.0 / (S(i, j) + S(i, jp)) * (Qp(i, j) - Qp(i, jp)) * invdxtheta);
}

// recalculate_viscosity, SourceEuler.cpp:205-223
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_recalc_viscosity(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr)
    derived_cell<SPLIT>(P, i, j, false, false);
}

// recalculate_derived_disk_quantities, SourceEuler.cpp:225-249
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_derived(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr)
    derived_cell<SPLIT>(P, i, j, true, true);
}

// locally isothermal sound speed and scale height from the initial setup
// (compute_sound_speed_normal / compute_scale_height_old, isothermal branches)
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_iso_profile(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr)
    const double r = P.Rmed[i];
    const double vK = sqrt(P.G * P.mstar / r);
    const double h = P.aspect_ratio * pow(r, P.flaring);
    const double c_s = h * vK;
    fld<SPLIT>(P, FCPT_SOUNDSPEED).set(i, j, c_s);
    const double inv_omega_kepler = 1.0 / omega_kepler(P, r);
    fld<SPLIT>(P, FCPT_SCALE_HEIGHT).set(i, j, c_s * inv_omega_kepler);
    const double sig = fld<SPLIT>(P, FCPT_SIGMA)(i, j);
    const double pr = sig * (c_s * c_s);
    fld<SPLIT>(P, FCPT_PRESSURE).set(i, j, pr);
    fld<SPLIT>(P, FCPT_TEMPERATURE).set(i, j, P.mu / P.Rgas * pr / sig);
}

// ---------------------------------------------------------------------------
// compute_viscous_stress_tensor, viscosity.cpp:139-254
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_stress(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(nu, FCPT_VISCOSITY);
    FLD(S, FCPT_SIGMA);
    const int jn = j == ns - 1 ? 0 : j + 1;
    const double vr0 = vr(i, j), vr1 = vr(i + 1, j), va0 = vaz(i, j), va1 = vaz(i, jn);
    const double divv = (vr1 * P.Rinf[i + 1] - vr0 * P.Rinf[i]) * P.InvDiffRsupRb[i] +
                        (va1 - va0) * P.invdphi * P.InvRmed[i];
    fld<SPLIT>(P, FCPT_DIV_V).set(i, j, divv);
    const double n = nu(i, j), s = S(i, j);
    const double drr = (vr1 - vr0) * P.InvDiffRsup[i];
    fld<SPLIT>(P, FCPT_TAU_R_R).set(i, j, 2.0 * n * s * (drr - 1.0 / 3.0 * divv));
    const double dpp = (va1 - va0) * P.invdphi * P.InvRmed[i] + 0.5 * (vr1 + vr0) * P.InvRmed[i];
    fld<SPLIT>(P, FCPT_TAU_PHI_PHI).set(i, j, 2.0 * n * s * (dpp - 1.0 / 3.0 * divv));
    if (i >= 1) {
        const int jp = j == 0 ? ns - 1 : j - 1;
        const double dvazirdr =
            (va0 * P.InvRmed[i] - vaz(i - 1, j) * P.InvRmed[i - 1]) * P.InvDiffRmed[i];
        const double dvrdphi = (vr0 - vr(i, jp)) * P.invdphi;
        const double drp = P.Rinf[i] * dvazirdr + dvrdphi * P.InvRinf[i];
        const double n4 = 0.25 * (n + nu(i - 1, j) + nu(i, jp) + nu(i - 1, jp));
        const double s4 = 0.25 * (s + S(i - 1, j) + S(i, jp) + S(i - 1, jp));
        fld<SPLIT>(P, FCPT_TAU_R_PHI).set(i, j, n4 * s4 * drp);
    }
}

// update_velocities_with_viscosity, viscosity.cpp:355-426
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_visc_update(KParams P)
{
    CELL_LOOP_HEAD(1, P.nr - 1)
    FLD(S, FCPT_SIGMA);
    FLD(trr, FCPT_TAU_R_R);
    FLD(tpp, FCPT_TAU_PHI_PHI);
    FLD(trp, FCPT_TAU_R_PHI);
    const double dt = P.s->dt;
    const int jp = j == 0 ? ns - 1 : j - 1;
    {
        FLD(vaz, FCPT_VAZI);
        const double sigma_avg = 0.5 * (S(i, j) + S(i, jp));
        const double a1 = P.Rinf[i + 1] * P.Rinf[i + 1], a0 = P.Rinf[i] * P.Rinf[i];
        const double dVp = dt * P.InvRmed[i] / (sigma_avg) *
                           ((2.0 / (a1 - a0)) * (a1 * trp(i + 1, j) - a0 * trp(i, j)) +
                            (tpp(i, j) - tpp(i, jp)) * P.invdphi);
        vaz.set(i, j, vaz(i, j) + dVp);
    }
    if (i >= 2) {
        FLD(vr, FCPT_VRAD);
        const int jn = j == ns - 1 ? 0 : j + 1;
        const double sigma_avg = 0.5 * (S(i, j) + S(i - 1, j));
        const double dVr = dt / (sigma_avg)*P.radial_viscosity_factor * 2.0 /
                           (P.Rmed[i] + P.Rmed[i - 1]) *
                           ((P.Rmed[i] * trr(i, j) - P.Rmed[i - 1] * trr(i - 1, j)) * P.InvDiffRmed[i] +
                            (trp(i, jn) - trp(i, j)) * P.invdphi - 0.5 * (tpp(i, j) + tpp(i - 1, j)));
        vr.set(i, j, vr(i, j) + dVr);
    }
}

// ---------------------------------------------------------------------------
// Q- (thermal_relaxation, SourceEuler.cpp:632-690) and Q+ (viscous_heating,
// SourceEuler.cpp:496-536) for one active cell
template <bool SPLIT>
__device__ __forceinline__ void heating_cooling_cell(const KParams &P, int i, int j, double time,
                                                     double e, double sig, double &qplus,
                                                     double &qminus)
{
    const int ns = P.ns;
    qminus = 0.0;
    if (P.cooling_beta_enabled) {
        const double r = P.Rmed[i];
        const double omega_k = omega_kepler(P, r);
        double beta_inv = 1 / P.cooling_beta;
        if (P.cooling_beta_ramp_up > 0.0) {
            const double x = 2 * time / P.cooling_beta_ramp_up;
            const double ramp_factor = 1 - exp(-(x * x));
            beta_inv = beta_inv * ramp_factor;
        }
        double delta_E = e;
        if (P.cooling_beta_ref == FCPT_BETA_REF_REFERENCE)
            delta_E -= fld<SPLIT>(P, FCPT_ENERGY0)(i, j) / fld<SPLIT>(P, FCPT_SIGMA0)(i, j) * sig;
        if (P.cooling_beta_ref == FCPT_BETA_REF_FLOOR)
            delta_E -= P.t_min * sig / P.mu * P.Rgas / (P.gamma - 1.0);
        qminus = 0.0 + delta_E * omega_k * beta_inv;
    }
    qplus = 0.0;
    if (P.heating_viscous) {
        const double nu = fld<SPLIT>(P, FCPT_VISCOSITY)(i, j);
        if (nu != 0.0) {
            FLD(trp, FCPT_TAU_R_PHI);
            const int jn = j == ns - 1 ? 0 : j + 1;
            const double tau_r_phi = 0.25 * (trp(i, j) + trp(i + 1, j) + trp(i, jn) + trp(i + 1, jn));
            const double trr = fld<SPLIT>(P, FCPT_TAU_R_R)(i, j);
            const double tpp = fld<SPLIT>(P, FCPT_TAU_PHI_PHI)(i, j);
            const double dv = fld<SPLIT>(P, FCPT_DIV_V)(i, j);
            double q = 1.0 / (2.0 * nu * sig) * (trr * trr + 2 * (tau_r_phi * tau_r_phi) + tpp * tpp);
            q += (2.0 / 9.0) * nu * sig * (dv * dv);
            q *= P.heating_viscous_factor;
            qplus = 0.0 + q;
        }
    }
}

__device__ __forceinline__ double rad_alpha(const KParams &P, double h, double sig, double e)
{
    const double inv_pow4 = pow(P.mu * (P.gamma - 1.0) / (P.Rgas * sig), 4.0);
    return 1.0 + 2.0 * h * 4.0 * P.sigma_sb / P.c_light * inv_pow4 * pow(e, 3.0);
}

// SubStep3, SourceEuler.cpp:859-954
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_substep3(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(E, FCPT_ENERGY);
    const double sig = fld<SPLIT>(P, FCPT_SIGMA)(i, j);
    double e = E(i, j);
    // compute_temperature (SourceEuler.cpp:1378-1408) before the update
    const double c_v_inv = P.mu / P.Rgas * (P.gamma - 1.0);
    fld<SPLIT>(P, FCPT_TEMPERATURE).set(i, j, c_v_inv * e / sig);
    double qplus = 0.0, qminus = 0.0;
    if (i >= 1 && i < P.nr - 1) {
        const DevScalars *s = P.s;
        heating_cooling_cell<SPLIT>(P, i, j, s->time, e, sig, qplus, qminus);
        const double alpha = rad_alpha(P, fld<SPLIT>(P, FCPT_SCALE_HEIGHT)(i, j), sig, e);
        qplus /= alpha;
        qminus /= alpha;
        double energy_new = e + s->dt * (qplus - qminus);
        const double SigmaFloor = 10.0 * P.sigma0 * P.sigma_floor;
        if ((sig < SigmaFloor)) {
            // TAU_EFF is never computed without surface cooling: it holds 0
            const double tau_eff = 0.0;
            const double e4 = qplus * tau_eff / (2.0 * P.sigma_sb);
            const double constant = (P.Rgas / P.mu * sig / (P.gamma - 1.0));
            energy_new = pow(e4, 1.0 / 4.0) * constant;
            qminus = qplus;
        }
        e = energy_new;
    }
    fld<SPLIT>(P, FCPT_QPLUS).set(i, j, qplus);
    fld<SPLIT>(P, FCPT_QMINUS).set(i, j, qminus);
    E.set(i, j, temperature_floor(P, e, sig));
}

// compute_heating_cooling_for_CFL, SourceEuler.cpp:1410-1450 (after update_viscosity
// and the stress tensor)
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_heatcool_cfl(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr)
    double qplus = 0.0, qminus = 0.0;
    if (i >= 1 && i < P.nr - 1) {
        const double e = fld<SPLIT>(P, FCPT_ENERGY)(i, j);
        const double sig = fld<SPLIT>(P, FCPT_SIGMA)(i, j);
        heating_cooling_cell<SPLIT>(P, i, j, P.s->time, e, sig, qplus, qminus);
        const double alpha = rad_alpha(P, fld<SPLIT>(P, FCPT_SCALE_HEIGHT)(i, j), sig, e);
        qplus /= alpha;
        qminus /= alpha;
    }
    fld<SPLIT>(P, FCPT_QPLUS).set(i, j, qplus);
    fld<SPLIT>(P, FCPT_QMINUS).set(i, j, qminus);
}

// ---------------------------------------------------------------------------
// damping (boundary_conditions/damping.cpp:311-427, 754-776): block row y maps
// to an inner-zone ring (y < n_inner) or an outer-zone ring
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_damping(KParams P, int n_inner)
{
    const int ns = P.ns;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (j >= ns) return;
    const double dt = P.s->dt;
    const double lin = P.damping_inner_limit, lout = P.damping_outer_limit;
    const double RMIN = P.rmin, RMAX = P.rmax;
    int i;
    bool inner;
    if (y < n_inner) {
        i = y;
        inner = true;
    } else {
        // outer zone starts at the smaller of the scalar / vector limits
        const int o0 = P.damp_ids[2] < P.damp_ids[3] ? P.damp_ids[2] : P.damp_ids[3];
        i = o0 + (y - n_inner);
        if (i > P.nr) return;
        inner = false;
    }
    const int F0[4] = {FCPT_VRAD0, FCPT_VAZI0, FCPT_SIGMA0, FCPT_ENERGY0};
    const int F[4] = {FCPT_VRAD, FCPT_VAZI, FCPT_SIGMA, FCPT_ENERGY};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const bool vector = q == 0;
        const double *radius = vector ? P.Rinf : P.Rmed;
        const int size_radial = vector ? P.nr + 1 : P.nr;
        if (i >= size_radial) continue;
        const Fld<SPLIT> X = fld<SPLIT>(P, F[q]), X0 = fld<SPLIT>(P, F0[q]);
        if (inner) {
            if (!((lin > 1.0) && (radius[0] < RMIN * lin))) continue;
            const int limit = vector ? P.damp_ids[1] : P.damp_ids[0];
            if (i > limit) continue;
            const double tau = P.damping_time_factor * 2.0 * M_PI / omega_kepler(P, RMIN);
            const double a = (radius[i] - RMIN * lin) / (RMIN - RMIN * lin);
            const double exp_factor = exp(-dt * (a * a) / tau);
            const double x = X(i, j), x0 = X0(i, j);
            X.set(i, j, (x - x0) * exp_factor + x0);
        } else {
            if (!((lout < 1.0) && (radius[size_radial - 1] > RMAX * lout))) continue;
            const int limit = vector ? P.damp_ids[3] : P.damp_ids[2];
            if (i < limit) continue;
            const double tau =
                P.damping_time_factor * 2.0 * M_PI / omega_kepler(P, P.damping_time_radius_outer);
            const double a = (radius[i] - RMAX * lout) / (RMAX - RMAX * lout);
            const double exp_factor = exp(-dt * (a * a) / tau);
            const double x = X(i, j), x0 = X0(i, j);
            X.set(i, j, (x - x0) * exp_factor + x0);
        }
    }
}

// ghost rings: zero-gradient Sigma/E, reflecting / zero-gradient / outflow v_r,
// Keplerian v_phi (boundary_conditions.cpp:67-97 and the BC files it dispatches to)
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_ghost_bc(KParams P, int advance_time)
{
    const int ns = P.ns;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (advance_time && j == 0) {
        // end of a sim::run iteration: time += dt and the monitor counter
        // (simulation.cpp:244, 545-551); nothing else in this kernel reads them
        DevScalars *s = P.s;
        const double time_next_monitor = (s->n_monitor + 1) * P.monitor_timestep;
        s->time = s->time + s->dt;
        if (advance_time == 2 && fabs(time_next_monitor - s->time) < 1e-6 * s->last_dt)
            s->n_monitor++;
    }
    if (j >= ns) return;
    const int nr = P.nr;
    FLD(S, FCPT_SIGMA);
    FLD(E, FCPT_ENERGY);
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    S.set(0, j, S(1, j));
    S.set(nr - 1, j, S(nr - 2, j));
    E.set(0, j, E(1, j));
    E.set(nr - 1, j, E(nr - 2, j));
    // rows 2 and nr-2 are never written here (nr >= 8)
    const double v2 = vr(2, j), vn2 = vr(nr - 2, j);
    if (P.bc_inner == FCPT_BC_REFLECTING) {
        vr.set(0, j, -v2);
        vr.set(1, j, 0);
    } else if (P.bc_inner == FCPT_BC_ZEROGRADIENT) {
        vr.set(0, j, v2);
        vr.set(1, j, v2);
    } else {
        if (v2 > 0.0) {
            vr.set(1, j, 0.0);
            vr.set(0, j, 0.0);
        } else {
            vr.set(1, j, v2);
            vr.set(0, j, v2);
        }
    }
    if (P.bc_outer == FCPT_BC_REFLECTING) {
        vr.set(nr, j, -vn2);
        vr.set(nr - 1, j, 0);
    } else if (P.bc_outer == FCPT_BC_ZEROGRADIENT) {
        vr.set(nr, j, vn2);
        vr.set(nr - 1, j, vn2);
    } else {
        if (vn2 < 0.0) {
            vr.set(nr - 1, j, 0.0);
            vr.set(nr, j, 0.0);
        } else {
            vr.set(nr - 1, j, vn2);
            vr.set(nr, j, vn2);
        }
    }
    const double vk_in = sqrt(P.G * P.mstar / P.Rmed[0]);
    const double vk_out = sqrt(P.G * P.mstar / P.Rmed[nr - 1]);
    vaz.set(0, j, 1.0 * vk_in - P.Rmed[0] * P.omega_frame);
    vaz.set(nr - 1, j, 1.0 * vk_out - P.Rmed[nr - 1] * P.omega_frame);
}

// ---------------------------------------------------------------------------
// Transport

// van Leer slope of a radial profile at ring k (compute_star_radial first loop,
// TransportEuler.cpp:355-371); arguments are rings k-1..k+1
__device__ __forceinline__ double dq_radial(const KParams &P, int k, double qm, double q0, double qp)
{
    if (k == 0 || k == P.nr - 1) return 0.0;
    const double dqm = (q0 - qm) * P.InvDiffRmed[k];
    const double dqp = (qp - q0) * P.InvDiffRmed[k + 1];
    return van_leer_lim(dqp, dqm);
}

// upwind interface value at ring interface i (compute_star_radial second loop,
// TransportEuler.cpp:376-405); w0..w3 = rings i-2..i+1 of the specific quantity
__device__ __forceinline__ double star_radial(const KParams &P, int i, double w0, double w1,
                                              double w2, double w3, double v, double dt)
{
    if (i <= 0 || i >= P.nr) return 0.0;
    if (v > 0.0) {
        const double dq = dq_radial(P, i - 1, w0, w1, w2);
        return w1 + (P.Rmed[i] - P.Rmed[i - 1] - v * dt) * 0.5 * dq;
    }
    const double dq = dq_radial(P, i, w1, w2, w3);
    return w2 - (P.Rmed[i + 1] - P.Rmed[i] + v * dt) * 0.5 * dq;
}

// One ring of the radial sweep's window: Sigma, the six advected quantities
// (four momenta, E, Sigma itself), their specific values Q / Sigma and v_r at
// the ring's inner interface.  The momenta are formed with the expressions of
// compute_momenta_from_velocities (TransportEuler.cpp:471-495).  Rings outside
// the grid hold Sigma = 1 and zero quantities (their interface values are
// never used: the slopes there are zero and DensityStar vanishes).
struct RingVals {
    double s;
    double q[6];
    double w[6];
    double v;
};

template <bool SPLIT>
__device__ __forceinline__ RingVals load_ring(const KParams &P, int r, int j, int jn)
{
    // loads go to a clamped ring and the results are selected, which keeps
    // the window in registers
    const bool in = r >= 0 && r < P.nr;
    const int rc = in ? r : 0;
    const int rv = r < 0 ? 0 : (r > P.nr ? P.nr : r);
    FLD(S, FCPT_SIGMA);
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    const double sig = S(rc, j);
    const double rr = P.Rmed[rc];
    const double OmegaF = P.omega_frame;
    const double v0 = vr(rc, j);
    double q[6];
    q[0] = sig * vr(rc + 1, j);
    q[1] = sig * v0;
    q[2] = sig * (vaz(rc, jn) + rr * OmegaF) * rr;
    q[3] = sig * (vaz(rc, j) + rr * OmegaF) * rr;
    q[4] = P.adiabatic ? fld<SPLIT>(P, FCPT_ENERGY)(rc, j) : 0.0;
    q[5] = sig;
    RingVals R;
    const double vv = vr(rv, j);
    R.v = (r >= 0 && r <= P.nr) ? vv : 0.0;
    R.s = in ? sig : 1.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        R.q[k] = in ? q[k] : 0.0;
        R.w[k] = in ? q[k] / sig : 0.0;
    }
    return R;
}

// OneWindRad, TransportEuler.cpp:138-168 and VanLeerRadial :545-619.  Every
// advected quantity of ring i is updated from rings i-2..i+2 of itself and of
// the density before the sweep (DENSITY_INT), so all of them are done in one
// pass that reads the state and writes separate outputs.  A thread owns one
// sector and walks `tile` consecutive rings outward with a window of four
// rings in registers: each ring is loaded once and each interface value
// (DensityStar and every quantity's QStar) is formed once.
template <bool SPLIT>
__global__ void __launch_bounds__(64) k_radial_sweep(KParams P, int tile)
{
    const int ns = P.ns, nr = P.nr;
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i0 = blockIdx.y * tile;
    if (j >= ns || i0 >= nr) return;
    const int jn = j == ns - 1 ? 0 : j + 1;
    const double dt = P.s->dt;
    const bool adi = P.adiabatic != 0;
    void *const outp[6] = {P.mom_b[0], P.mom_b[1], P.mom_b[2], P.mom_b[3], P.en_b, P.sig_b};
    const double *const outbg[6] = {P.bgm[0], P.bgm[1], P.bgm[2], P.bgm[3], P.bg[FCPT_ENERGY],
                                    P.bg[FCPT_SIGMA]};
    // window at interface k: rings k-2 .. k+1
    RingVals R0 = load_ring<SPLIT>(P, i0 - 2, j, jn);
    RingVals R1 = load_ring<SPLIT>(P, i0 - 1, j, jn);
    RingVals R2 = load_ring<SPLIT>(P, i0, j, jn);
    RingVals R3 = load_ring<SPLIT>(P, i0 + 1, j, jn);
    double qs_prev[6], ds_prev = 0.0, v_prev = 0.0;
    const int iend = i0 + tile < nr ? i0 + tile : nr;
    for (int k = i0;; ++k) {
        const double v = R2.v;
        const double ds = star_radial(P, k, R0.s, R1.s, R2.s, R3.s, v, dt);
        double qs[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) qs[q] = star_radial(P, k, R0.w[q], R1.w[q], R2.w[q], R3.w[q], v, dt);
        if (k > i0) {
            // ring k-1 lies between interfaces k-1 and k
            const int i = k - 1;
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                if (q == 4 && !adi) continue;
                const double varq_inf = dt * P.dphi * P.Rinf[i] * qs_prev[q] * ds_prev * v_prev;
                const double varq_sup = dt * P.dphi * P.Rsup[i] * qs[q] * ds * v;
                buf<SPLIT>(P, outp[q], outbg[q]).set(i, j, R1.q[q] + (varq_inf - varq_sup) * P.InvSurf[i]);
            }
        }
        if (k == iend) break;
#pragma unroll
        for (int q = 0; q < 6; ++q) qs_prev[q] = qs[q];
        ds_prev = ds;
        v_prev = v;
        R0 = R1;
        R1 = R2;
        R2 = R3;
        R3 = load_ring<SPLIT>(P, k + 2, j, jn);
    }
}

// ComputeStarTheta for one interface j (TransportEuler.cpp:416-465);
// q[0..3] = sectors j-2..j+1
__device__ __forceinline__ double star_theta(const double *q, double dxtheta, double invdxtheta,
                                             double ksi)
{
    if (ksi > 0.0) {
        const double dq = 0.5 * van_leer_lim(q[2] - q[1], q[1] - q[0]) * invdxtheta;
        return q[1] + (dxtheta - ksi) * dq;
    }
    const double dq = 0.5 * van_leer_lim(q[3] - q[2], q[2] - q[1]) * invdxtheta;
    return q[2] - (dxtheta + ksi) * dq;
}

// OneWindTheta, TransportEuler.cpp:270-289, in one pass per ring segment:
// QuantitiesAdvection (:292-304, VanLeerTheta :630-664) with the residual
// velocity v - <v>, then with the uniform residual of the ring, then
// AdvectSHIFT (:238-268).  A block owns kAzSeg sectors of one ring and stages
// them with a halo of 4 sectors in LDS (2 per van Leer pass); the result of
// sector j is written to sector j + Nshift.  The density is advanced first
// because the second pass needs its first-pass value as DENSITY_INT; the other
// quantities only read the density, so their order does not matter.
constexpr int kAzSeg = 256;
constexpr int kAzHalo = 4;
constexpr int kAzExt = kAzSeg + 2 * kAzHalo;

template <bool SPLIT>
__global__ void __launch_bounds__(256) k_azimuthal_sweep(KParams P)
{
    __shared__ double sS0[kAzExt], sV0[kAzExt], sDS0[kAzExt], sS1[kAzExt], sDS1[kAzExt];
    __shared__ double sQ0[kAzExt], sW[kAzExt], sQS[kAzExt], sQ1[kAzExt];
    const int ns = P.ns;
    const int i = blockIdx.y;
    const int j0 = blockIdx.x * kAzSeg;
    const int Ls = ns - j0 < kAzSeg ? ns - j0 : kAzSeg;
    const int next = Ls + 2 * kAzHalo;  // staged positions p: sector j0 - 4 + p
    const int t = threadIdx.x;
    const double dt = P.s->dt;
    const double vm = P.vmean[i];
    const double vu = P.vuni[i];
    const int nsh = P.nshift[i];
    const double dxtheta = P.dphi * P.Rmed[i];
    const double invdxtheta = 1.0 / dxtheta;
    const double dxrad = (P.Rsup[i] - P.Rinf[i]) * dt;
    const double invsurf = P.InvSurf[i];
    auto sector = [ns](int j) {
        if (j < 0) j += ns;
        if (j >= ns) j -= ns;
        return j;
    };
    auto dst_sector = [ns](int j) {
        while (j < 0) j += ns;
        while (j >= ns) j -= ns;
        return j;
    };
    FLD(vaz, FCPT_VAZI);
    const Fld<SPLIT> Sin = buf<SPLIT>(P, P.sig_b, P.bg[FCPT_SIGMA]);
    for (int p = t; p < next; p += 256) {
        const int j = sector(j0 - kAzHalo + p);
        sS0[p] = Sin(i, j);
        sV0[p] = vaz(i, j) - vm;
    }
    __syncthreads();
    // first pass (residual velocity): DensityStar on interfaces 2 .. next-2
    for (int p = 2 + t; p <= next - 2; p += 256)
        sDS0[p] = star_theta(&sS0[p - 2], dxtheta, invdxtheta, sV0[p] * dt);
    // the density's own first pass
    for (int p = t; p < next; p += 256) sW[p] = sS0[p] / sS0[p];
    __syncthreads();
    for (int p = 2 + t; p <= next - 2; p += 256)
        sQS[p] = star_theta(&sW[p - 2], dxtheta, invdxtheta, sV0[p] * dt);
    __syncthreads();
    for (int p = 2 + t; p < next - 2; p += 256) {
        double varq = dxrad * sQS[p] * sDS0[p] * sV0[p];
        varq -= dxrad * sQS[p + 1] * sDS0[p + 1] * sV0[p + 1];
        sS1[p] = sS0[p] + varq * invsurf;
    }
    __syncthreads();
    // second pass (uniform residual): DensityStar on interfaces 4 .. 4+Ls
    for (int p = kAzHalo + t; p <= kAzHalo + Ls; p += 256)
        sDS1[p] = star_theta(&sS1[p - 2], dxtheta, invdxtheta, vu * dt);
    for (int p = 2 + t; p < next - 2; p += 256) sW[p] = sS1[p] / sS1[p];
    __syncthreads();
    for (int p = kAzHalo + t; p <= kAzHalo + Ls; p += 256)
        sQS[p] = star_theta(&sW[p - 2], dxtheta, invdxtheta, vu * dt);
    __syncthreads();
    {
        const Fld<SPLIT> Sout = buf<SPLIT>(P, P.sig_c, P.bg[FCPT_SIGMA]);
        for (int p = kAzHalo + t; p < kAzHalo + Ls; p += 256) {
            double varq = dxrad * sQS[p] * sDS1[p] * vu;
            varq -= dxrad * sQS[p + 1] * sDS1[p + 1] * vu;
            Sout.set(i, dst_sector(j0 - kAzHalo + p + nsh), sS1[p] + varq * invsurf);
        }
    }
    // the other quantities
    const int nq = P.adiabatic ? 5 : 4;
    for (int q = 0; q < nq; ++q) {
        const Fld<SPLIT> Qin = q < 4 ? buf<SPLIT>(P, P.mom_b[q], P.bgm[q])
                                     : buf<SPLIT>(P, P.en_b, P.bg[FCPT_ENERGY]);
        const Fld<SPLIT> out = q < 4 ? buf<SPLIT>(P, P.mom[q], P.bgm[q])
                                     : buf<SPLIT>(P, P.en_c, P.bg[FCPT_ENERGY]);
        __syncthreads();
        for (int p = t; p < next; p += 256) {
            const double Qv = Qin(i, sector(j0 - kAzHalo + p));
            sQ0[p] = Qv;
            sW[p] = Qv / sS0[p];
        }
        __syncthreads();
        for (int p = 2 + t; p <= next - 2; p += 256)
            sQS[p] = star_theta(&sW[p - 2], dxtheta, invdxtheta, sV0[p] * dt);
        __syncthreads();
        for (int p = 2 + t; p < next - 2; p += 256) {
            double varq = dxrad * sQS[p] * sDS0[p] * sV0[p];
            varq -= dxrad * sQS[p + 1] * sDS0[p + 1] * sV0[p + 1];
            const double q1 = sQ0[p] + varq * invsurf;
            sQ1[p] = q1;
            sW[p] = q1 / sS1[p];
        }
        __syncthreads();
        for (int p = kAzHalo + t; p <= kAzHalo + Ls; p += 256)
            sQS[p] = star_theta(&sW[p - 2], dxtheta, invdxtheta, vu * dt);
        __syncthreads();
        for (int p = kAzHalo + t; p < kAzHalo + Ls; p += 256) {
            double varq = dxrad * sQS[p] * sDS1[p] * vu;
            varq -= dxrad * sQS[p + 1] * sDS1[p + 1] * vu;
            out.set(i, dst_sector(j0 - kAzHalo + p + nsh), sQ1[p] + varq * invsurf);
        }
    }
}

// compute_velocities_from_momenta (TransportEuler.cpp:498-535), then
// assure_minimum_value(Sigma) (SourceEuler.cpp:102-131) and, adiabatic,
// SetTemperatureFloorCeilValues (TransportEuler.cpp:127-131)
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_velocities(KParams P)
{
    CELL_LOOP_HEAD(0, P.nr)
    const Fld<SPLIT> Sb = buf<SPLIT>(P, P.sig_c, P.bg[FCPT_SIGMA]);
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    const int jm = j == 0 ? ns - 1 : j - 1;
    const double s0 = Sb(i, j);
    if (i == 0) {
        vr.set(i, j, 0.0);
    } else {
        const Fld<SPLIT> m0 = buf<SPLIT>(P, P.mom[0], P.bgm[0]);
        const Fld<SPLIT> m1 = buf<SPLIT>(P, P.mom[1], P.bgm[1]);
        vr.set(i, j, (m0(i - 1, j) + m1(i, j)) / (Sb(i - 1, j) + s0));
    }
    const Fld<SPLIT> m2 = buf<SPLIT>(P, P.mom[2], P.bgm[2]);
    const Fld<SPLIT> m3 = buf<SPLIT>(P, P.mom[3], P.bgm[3]);
    vaz.set(i, j, (m2(i, jm) + m3(i, j)) / (Sb(i, jm) + s0) * P.InvRmed[i] - P.Rmed[i] * P.omega_frame);
    const double floor_value = P.sigma_floor * P.sigma0;
    double sig = s0;
    if (sig < floor_value) sig = floor_value;
    fld<SPLIT>(P, FCPT_SIGMA).set(i, j, sig);
    if (P.adiabatic) {
        const Fld<SPLIT> Eb = buf<SPLIT>(P, P.en_c, P.bg[FCPT_ENERGY]);
        fld<SPLIT>(P, FCPT_ENERGY).set(i, j, temperature_floor(P, Eb(i, j), sig));
    }
}

// ---------------------------------------------------------------------------
// totals with Neumaier compensation: per ring (sum, compensation) pairs, then
// one block folds the rings in order (quantities.cpp:51-78, 242-279)
__device__ __forceinline__ void neumaier_add(double &s, double &c, double x)
{
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}

template <bool SPLIT>
__global__ void __launch_bounds__(64) k_totals_rings(KParams P, double radius, double *ring_out)
{
    const int ns = P.ns;
    const int i = blockIdx.x;  // ring
    const int t = threadIdx.x;
    double m = 0.0, mc = 0.0, l = 0.0, lc = 0.0;
    const bool active = i >= 1 && i < P.nr - 1 && P.Rmed[i] <= radius;
    if (active) {
        FLD(S, FCPT_SIGMA);
        FLD(vaz, FCPT_VAZI);
        for (int j = t; j < ns; j += 64) {
            const int jm = j == 0 ? ns - 1 : j - 1;
            const double s = S(i, j);
            neumaier_add(m, mc, P.Surf[i] * s);
            neumaier_add(l, lc, P.Surf[i] * 0.5 * (s + S(i, jm)) * P.Rmed[i] *
                                    (vaz(i, j) + P.omega_frame * P.Rmed[i]));
        }
    }
    __shared__ double sh[4][64];
    sh[0][t] = m;
    sh[1][t] = mc;
    sh[2][t] = l;
    sh[3][t] = lc;
    __syncthreads();
    if (t == 0) {
        double M = 0.0, MC = 0.0, L = 0.0, LC = 0.0;
        for (int k = 0; k < 64; ++k) {
            neumaier_add(M, MC, sh[0][k]);
            MC += sh[1][k];
            neumaier_add(L, LC, sh[2][k]);
            LC += sh[3][k];
        }
        ring_out[4 * i + 0] = M;
        ring_out[4 * i + 1] = MC;
        ring_out[4 * i + 2] = L;
        ring_out[4 * i + 3] = LC;
    }
}

__global__ void k_totals_final(KParams P, const double *ring_out, double *out2)
{
    double M = 0.0, MC = 0.0, L = 0.0, LC = 0.0;
    for (int i = 0; i < P.nr; ++i) {
        neumaier_add(M, MC, ring_out[4 * i + 0]);
        MC += ring_out[4 * i + 1];
        neumaier_add(L, LC, ring_out[4 * i + 2]);
        LC += ring_out[4 * i + 3];
    }
    out2[0] = M + MC;
    out2[1] = L + LC;
}

// streaming copy of four work arrays (bandwidth probe), 16-byte accesses
__global__ void __launch_bounds__(256) k_copy4(KParams P, size_t n16)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (int q = 0; q < 4; ++q) {
        const double2 *src = static_cast<const double2 *>(P.mom[q]);
        double2 *dst = static_cast<double2 *>(P.mom_b[q]);
        for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n16; k += stride)
            dst[k] = src[k];
    }
}

// split storage: re-express a field's deviations about a new ring background
__global__ void __launch_bounds__(256) k_rebase(float *d, const double *oldbg, const double *newbg,
                                                int ns, int rows)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (j >= ns || i >= rows) return;
    const size_t c = IX(i, j);
    d[c] = (float)((oldbg[i] + (double)d[c]) - newbg[i]);
}

// ---------------------------------------------------------------------------
// launch helpers: each dispatches on the storage mode

#define LAUNCH(kern, grid, block, st, ...)                                               \
    do {                                                                                 \
        if (P.split) hipLaunchKernelGGL(kern<true>, grid, block, 0, st, __VA_ARGS__);    \
        else hipLaunchKernelGGL(kern<false>, grid, block, 0, st, __VA_ARGS__);           \
    } while (0)

static inline dim3 cell_grid(const KParams &P, int rows)
{
    return dim3((unsigned)((P.ns + 63) / 64), (unsigned)((rows + 3) / 4));
}
static const dim3 kBlock(64, 4);

size_t element_size(const KParams &P) { return P.split ? sizeof(float) : sizeof(double); }

void launch_ring_mean(const KParams &P, hipStream_t st, int mode)
{
    LAUNCH(k_ring_mean, dim3(P.nr), dim3(256), st, P, mode);
}

// CFL grid: at most kCflMaxBlocks blocks, rows walked with a grid stride
static dim3 cfl_grid(const KParams &P)
{
    const unsigned gx = (unsigned)((P.ns + 63) / 64);
    unsigned gy = (unsigned)((P.nr - 2 + 3) / 4);
    unsigned cap = kCflMaxBlocks / gx;
    if (cap < 1) cap = 1;
    if (gy > cap) gy = cap;
    if (gy < 1) gy = 1;
    return dim3(gx, gy);
}

void launch_cfl(const KParams &P, hipStream_t st, int mode)
{
    launch_ring_mean(P, st, 0);
    const dim3 g = cfl_grid(P);
    LAUNCH(k_cfl_cells, g, kBlock, st, P);
    hipLaunchKernelGGL(k_cfl_finalize, dim3(1), dim3(256), 0, st, P, mode, (int)(g.x * g.y));
}

void launch_cfl_cells(const KParams &P, hipStream_t st) { LAUNCH(k_cfl_cells, cfl_grid(P), kBlock, st, P); }

void launch_set_dt(const KParams &P, hipStream_t st, double dt)
{
    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, st, P, dt);
}
void launch_set_clock(const KParams &P, hipStream_t st, double t)
{
    hipLaunchKernelGGL(k_set_clock, dim3(1), dim3(1), 0, st, P, t);
}
void launch_set_time(const KParams &P, hipStream_t st, double t, double last_dt, int n_monitor)
{
    hipLaunchKernelGGL(k_set_time, dim3(1), dim3(1), 0, st, P, t, last_dt, n_monitor);
}
void launch_set_params(const KParams &P, hipStream_t st, const fcpt_step_params &sp)
{
    hipLaunchKernelGGL(k_set_params, dim3(1), dim3(1), 0, st, P, sp);
}

void launch_potential(const KParams &P, hipStream_t st) { LAUNCH(k_potential, cell_grid(P, P.nr), kBlock, st, P); }

void launch_sources(const KParams &P, hipStream_t st)
{
    LAUNCH(k_src_vr, cell_grid(P, P.nr - 3), kBlock, st, P);
    LAUNCH(k_src_vphi, cell_grid(P, P.nr - 2), kBlock, st, P);
    if (P.adiabatic) LAUNCH(k_src_energy, cell_grid(P, P.nr - 1), kBlock, st, P);
}

// returns true when the viscosity recalculation was folded in
bool launch_artificial_viscosity(const KParams &P, hipStream_t st, int with_viscosity)
{
    if (P.artvisc == FCPT_ARTVISC_TW) {
        LAUNCH(k_tw_q, cell_grid(P, P.nr), kBlock, st, P);
        LAUNCH(k_tw_vel, cell_grid(P, P.nr), kBlock, st, P, with_viscosity);
        return with_viscosity != 0;
    }
    if (P.artvisc == FCPT_ARTVISC_SN) {
        LAUNCH(k_sn_q, cell_grid(P, P.nr), kBlock, st, P);
        LAUNCH(k_sn_vel, cell_grid(P, P.nr), kBlock, st, P, with_viscosity);
        return with_viscosity != 0;
    }
    if (P.adiabatic && P.artvisc_dissipation) LAUNCH(k_temperature_floor, cell_grid(P, P.nr), kBlock, st, P);
    return false;
}

void launch_recalc_viscosity(const KParams &P, hipStream_t st)
{
    LAUNCH(k_recalc_viscosity, cell_grid(P, P.nr), kBlock, st, P);
}
void launch_stress(const KParams &P, hipStream_t st) { LAUNCH(k_stress, cell_grid(P, P.nr), kBlock, st, P); }
void launch_visc_update(const KParams &P, hipStream_t st)
{
    LAUNCH(k_visc_update, cell_grid(P, P.nr - 2), kBlock, st, P);
}
void launch_substep3(const KParams &P, hipStream_t st) { LAUNCH(k_substep3, cell_grid(P, P.nr), kBlock, st, P); }
void launch_heatcool_cfl(const KParams &P, hipStream_t st)
{
    LAUNCH(k_heatcool_cfl, cell_grid(P, P.nr), kBlock, st, P);
}

void launch_boundaries(const KParams &P, hipStream_t st, int final_call, int advance_time)
{
    if (final_call && P.damping) {
        const int n_inner = (P.damp_ids[0] > P.damp_ids[1] ? P.damp_ids[0] : P.damp_ids[1]) + 1;
        const int o0 = P.damp_ids[2] < P.damp_ids[3] ? P.damp_ids[2] : P.damp_ids[3];
        const int n_outer = P.nr + 1 - o0;
        LAUNCH(k_damping, cell_grid(P, n_inner + n_outer), kBlock, st, P, n_inner);
    }
    LAUNCH(k_ghost_bc, dim3((unsigned)((P.ns + 255) / 256)), dim3(256), st, P, advance_time);
}

// rings per radial-sweep thread: the longest walk that still gives every SIMD
// of the 256 CUs three wavefronts (the kernel's register budget allows about
// three), so that the four-ring window start is amortised where the grid is
// large and the chip stays full where it is small
int radial_tile_for(int nr, int ns, int requested)
{
    if (requested > 0) return requested;
    const long waves_per_band = (long)((ns + 63) / 64);
    int tile = 32;
    while (tile > 2 && waves_per_band * ((nr + tile - 1) / tile) < 3072) tile /= 2;
    return tile;
}

void launch_radial_sweep(const KParams &P, hipStream_t st)
{
    const int tile = P.radial_tile > 0 ? P.radial_tile : radial_tile_for(P.nr, P.ns, 0);
    const dim3 g((unsigned)((P.ns + 63) / 64), (unsigned)((P.nr + tile - 1) / tile));
    LAUNCH(k_radial_sweep, g, dim3(64), st, P, tile);
}
void launch_azimuthal_sweep(const KParams &P, hipStream_t st)
{
    const dim3 g((unsigned)((P.ns + kAzSeg - 1) / kAzSeg), (unsigned)P.nr);
    LAUNCH(k_azimuthal_sweep, g, dim3(256), st, P);
}

void launch_copy4(const KParams &P, hipStream_t st)
{
    const size_t n16 = (size_t)P.nr * P.ns * element_size(P) / 16;
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_copy4, dim3((unsigned)blocks), dim3(256), 0, st, P, n16);
}

void launch_transport(const KParams &P, hipStream_t st)
{
    launch_radial_sweep(P, st);
    launch_ring_mean(P, st, 1);
    launch_azimuthal_sweep(P, st);
    LAUNCH(k_velocities, cell_grid(P, P.nr), kBlock, st, P);
}

void launch_derived(const KParams &P, hipStream_t st) { LAUNCH(k_derived, cell_grid(P, P.nr), kBlock, st, P); }
void launch_iso_profile(const KParams &P, hipStream_t st)
{
    LAUNCH(k_iso_profile, cell_grid(P, P.nr), kBlock, st, P);
}

void launch_rebase(const KParams &P, hipStream_t st, void *field, const double *oldbg,
                   const double *newbg, int rows)
{
    hipLaunchKernelGGL(k_rebase, cell_grid(P, rows), kBlock, 0, st, static_cast<float *>(field), oldbg,
                       newbg, P.ns, rows);
}

void launch_totals(const KParams &P, hipStream_t st, double radius, double *ring_scratch,
                   double *out2)
{
    LAUNCH(k_totals_rings, dim3(P.nr), dim3(64), st, P, radius, ring_scratch);
    hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(1), 0, st, P, ring_scratch, out2);
}

}  // namespace fcpt
