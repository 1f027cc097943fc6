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
#include <vector>

#include "fcpt_device.hpp"
#include "fcpt_kernels.hpp"

namespace fcpt {

#define IX(i, j) ((size_t)(i) * (size_t)ns + (size_t)(j))

__device__ __forceinline__ double std_min(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double std_max(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double van_leer_lim(double a, double b)
{
    // TransportEuler.cpp:306-312, evaluated without a branch (the quotient of
    // a rejected pair is discarded)
    const double r = 2.0 * a * b / (a + b);
    return (a * b > 0.0) ? r : 0.0;
}
// Split storage rounds every stored value to an FP32 deviation, so its
// kernels may form quantities that only scale a correction term in FP32 and
// replace divisions by a reciprocal; the FP64 instantiations keep the
// reference's IEEE operations.  van Leer limiter: FP32 in split storage (the
// slope multiplies the second-order term of the interface value), its
// quotient through the hardware reciprocal (1 ulp; an FP32 division without
// fast-math is the IEEE sequence of about ten instructions).
template <bool SPLIT>
__device__ __forceinline__ double vl_lim(double a, double b)
{
    if constexpr (SPLIT) {
        const float fa = (float)a, fb = (float)b;
        const float r = 2.0f * fa * fb * __builtin_amdgcn_rcpf(fa + fb);
        return (fa * fb > 0.0f) ? (double)r : 0.0;
    } else {
        return van_leer_lim(a, b);
    }
}

// Quotients n / d that share one denominator d.  The compiler's IEEE FP64
// division on gfx950 is v_div_scale (range scaling of d and n), v_rcp, two
// Newton steps on the reciprocal, the quotient, one FMA correction
// (v_div_fmas) and v_div_fixup (zeros, infinities, NaNs, out-of-range
// results).  For |d| in [2^-250, 2^250) and |n| in [2^-400, 2^250) neither
// the scaling nor the fix-up changes anything, so the same reciprocal and
// correction steps written out give the same bits, and the reciprocal part
// (five instructions) is formed once per denominator: a quotient then costs
// three.  Other operands (zero aside, which returns the signed zero n * r)
// take the division itself.
struct SharedDen {
    double d, r;
    bool ok;
};
__device__ __forceinline__ SharedDen shared_den(double d)
{
    const double r0 = __builtin_amdgcn_rcp(d);
    const double e0 = fma(-d, r0, 1.0);
    const double r1 = fma(r0, e0, r0);
    const double e1 = fma(-d, r1, 1.0);
    const double r2 = fma(r1, e1, r1);
    const unsigned ed = (unsigned)((unsigned long long)__double_as_longlong(d) >> 52) & 0x7ffu;
    return SharedDen{d, r2, ed - 773u < 500u};
}
__device__ __forceinline__ double shared_div(double n, const SharedDen &D)
{
    const double q0 = n * D.r;
    const double rem = fma(-D.d, q0, n);
    double q = fma(rem, D.r, q0);
    const unsigned en = (unsigned)((unsigned long long)__double_as_longlong(n) >> 52) & 0x7ffu;
    if (n == 0.0) q = q0;
    else if (!D.ok || en - 623u >= 650u) q = n / D.d;
    return q;
}

// IEEE FP64 square root.  The compiler's sqrt on gfx950 scales x by 2^256
// below 2^-767, refines v_rsq by one Goldschmidt step and two Newton
// corrections, scales back, and passes zeros and +inf through a class test.
// For x in [2^-767, inf) the scalings are by 2^0 and the class test passes
// the refined value, so the refinement written out gives the same bits;
// other operands take sqrt itself.
__device__ __forceinline__ double sqrt_ieee(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double e = fma(-g, g, x);
    g = fma(e, h, g);
    e = fma(-g, g, x);
    g = fma(e, h, g);
    if (!(x >= 0x1p-767 && x < __builtin_inf())) {
        __builtin_amdgcn_s_sleep(0);  // a branch: not both roots formed and selected
        g = sqrt(x);
    }
    return g;
}

// FP64 reciprocal: one IEEE division, or in split storage the hardware
// reciprocal refined by two Newton steps (error a few ulp)
template <bool SPLIT>
__device__ __forceinline__ double recip(double x)
{
    if constexpr (SPLIT) {
        double y = __builtin_amdgcn_rcp(x);
        y = y * (2.0 - x * y);
        return y * (2.0 - x * y);
    } else {
        return 1.0 / x;
    }
}

// 1 / sqrt(x): IEEE square root and division, or in split storage the
// hardware reciprocal square root refined by two Newton steps
template <bool SPLIT>
__device__ __forceinline__ double inv_sqrt(double x)
{
    if constexpr (SPLIT) {
        double y = __builtin_amdgcn_rsq(x);
        y = y * (1.5 - 0.5 * x * y * y);
        return y * (1.5 - 0.5 * x * y * y);
    } else {
        return 1.0 / sqrt(x);
    }
}

// n / d: the IEEE quotient, or in split storage the product of n with the
// refined reciprocal of d
template <bool SPLIT>
__device__ __forceinline__ double quot(double n, double d)
{
    if constexpr (SPLIT) return n * recip<true>(d);
    else return n / d;
}

__device__ __forceinline__ double omega_kepler(const KParams &P, double r)
{
    return sqrt(P.G * P.mstar / (r * r * r));  // Theo.cpp:246-249
}


// Per-ring arrays (metric, backgrounds, ring means, shifts) are never written
// by the kernel that reads them.  Reading them through the constant address
// space says so: with a wave-uniform index they become scalar loads, which
// use their own counter and do not wait behind the wavefront's vector loads.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(4))) T *cst(const T *p)
{
    return (const __attribute__((address_space(4))) T *)p;
}

// Split storage rounds each deviation to FP32 either to nearest or
// stochastically: up or down with probabilities proportional to the distances
// to the two neighbouring FP32 values (the rounding is unbiased, so the
// per-step roundings of a slowly changing value do not add up in one
// direction: an increment below half an FP32 unit of the deviation survives
// on average instead of being lost every step).  The uniform variate of a
// store is one multiply-xorshift round (Wellons' hash prospector, public
// domain) of the cell index, the deviation's high word (sign, exponent and
// leading mantissa bits: independent of the dropped bits, and different for
// the several values one field of one cell takes in a step) and a key drawn
// per field and step (sr_key, Wellons' lowbias32): a value always rounds the
// same way at its cell within one kernel (a thread can form what another
// thread stores) and differently from step to step.
__device__ __forceinline__ unsigned sr_hash(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
// one multiply-xorshift round (the first half of lowbias32)
__device__ __forceinline__ unsigned sr_bits(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    return x;
}
__device__ __forceinline__ float round_dev(double d, unsigned key, size_t c)
{
    if (!key) return (float)d;
    // add a 29-bit variate below the FP32 unit of the FP64 deviation and drop
    // those 29 bits: rounded away from zero with probability (dropped
    // bits) / 2^29.  Exact for |d| in [2^-126, 2^128) (the sign bit is never
    // reached); smaller deviations are rounded once more by the conversion,
    // infinities stay infinite.
    const unsigned long long b0 = (unsigned long long)__double_as_longlong(d);
    const unsigned u = sr_bits((unsigned)c ^ key ^ (unsigned)(b0 >> 32)) >> 3;
    const unsigned long long b = b0 + (unsigned long long)u;
    return (float)__longlong_as_double((long long)(b & ~0x1fffffffull));
}
__device__ __forceinline__ unsigned sr_key(int id, unsigned step)
{
    return sr_hash(step * 0x85ebca77u + (unsigned)id * 0xc2b2ae3du) | 1u;
}

// field accessor: an FP64 array, or FP32 deviations from an FP64 ring background
template <bool SPLIT>
struct Fld {
    using T = typename std::conditional<SPLIT, float, double>::type;
    T *p;
    const double *bg;
    int ns;
    unsigned key;  // split storage: 0 rounds to nearest, else the stochastic rounding's key
    __device__ __forceinline__ double operator()(int i, int j) const
    {
        const size_t c = (size_t)i * (size_t)ns + (size_t)j;
        if constexpr (SPLIT) return cst(bg)[i] + (double)p[c];
        else return p[c];
    }
    __device__ __forceinline__ void set(int i, int j, double v) const
    {
        const size_t c = (size_t)i * (size_t)ns + (size_t)j;
        if constexpr (SPLIT) p[c] = round_dev(v - cst(bg)[i], key, c);
        else p[c] = v;
    }
    // the stored value alone, and its decoding: lets a kernel issue a load
    // early and pay for it where the value is used
    __device__ __forceinline__ T raw(int i, int j) const { return p[(size_t)i * (size_t)ns + (size_t)j]; }
    __device__ __forceinline__ void set_raw(int i, int j, T x) const { p[(size_t)i * (size_t)ns + (size_t)j] = x; }
    // store v at (i, j) and return it as a later load of it returns it
    __device__ __forceinline__ double set_get(int i, int j, double v) const
    {
        const size_t c = (size_t)i * (size_t)ns + (size_t)j;
        if constexpr (SPLIT) {
            const double b = cst(bg)[i];
            const float x = round_dev(v - b, key, c);
            p[c] = x;
            return b + (double)x;
        } else {
            p[c] = v;
            return v;
        }
    }
    // v as a later load of the value stored at (i, j) returns it
    __device__ __forceinline__ double stored(int i, int j, double v) const
    {
        const size_t c = (size_t)i * (size_t)ns + (size_t)j;
        if constexpr (SPLIT) return cst(bg)[i] + (double)round_dev(v - cst(bg)[i], key, c);
        else return v;
    }
    __device__ __forceinline__ double dec(int i, T x) const
    {
        if constexpr (SPLIT) return cst(bg)[i] + (double)x;
        else return x;
    }
};

// a ring-uniform quotient of ring i (row t of KParams::ring_tab, formed once
// by k_ring_tables with the expression its users had inline: the same bits)
__device__ __forceinline__ double ring_tab(const KParams &P, int t, int i)
{
    return cst(P.ring_tab)[(size_t)t * (size_t)(P.nr + 2) + (size_t)i];
}

// the rounding key of split-storage stores of field (or buffer) `id` in the
// current step; 0 (to nearest) in FP64 storage, with stochastic rounding off,
// and (split_rounding 1) for the fields derived afresh from the state every
// step, whose roundings do not accumulate: stochastic rounding then applies
// to the evolved fields Sigma, v_r, v_phi, E and the transported momenta
constexpr int kBufMom = 64;
__device__ __forceinline__ bool carries_state(int id)
{
    return id == FCPT_SIGMA || id == FCPT_VRAD || id == FCPT_VAZI || id == FCPT_ENERGY || id >= kBufMom;
}
template <bool SPLIT>
__device__ __forceinline__ unsigned store_key(const KParams &P, int id)
{
    if constexpr (SPLIT) {
        const int mode = P.split_rounding;
        return (mode == 2 || (mode == 1 && carries_state(id))) ? sr_key(id, cst(&P.s->step_count)[0]) : 0u;
    } else {
        return 0u;
    }
}
template <bool SPLIT>
__device__ __forceinline__ Fld<SPLIT> fld(const KParams &P, int f)
{
    return Fld<SPLIT>{static_cast<typename Fld<SPLIT>::T *>(P.f[f]), P.bg[f], P.ns, store_key<SPLIT>(P, f)};
}
// a work buffer; id names it for the rounding key (the momenta: kBufMom + q)
template <bool SPLIT>
__device__ __forceinline__ Fld<SPLIT> buf(const KParams &P, void *p, const double *bg, int id = -1)
{
    return Fld<SPLIT>{static_cast<typename Fld<SPLIT>::T *>(p), bg, P.ns, id >= 0 ? store_key<SPLIT>(P, id) : 0u};
}

#define FLD(name, id) const Fld<SPLIT> name = fld<SPLIT>(P, id)

// The transported momenta after the azimuthal sweep, read by the end of the
// step: FP64 in both storage modes (in split storage their FP32 rounding,
// once per step, was the largest part of the error near the planets, and
// the sweep that stores them is bound by its instructions, which the
// stochastic rounding of four stores per cell adds to)
__device__ __forceinline__ Fld<false> mom_out(const KParams &P, int q)
{
    return Fld<false>{static_cast<double *>(P.mom[q]), nullptr, P.ns, 0u};
}

// The gas potential is stored in FP64 in both storage modes: its radial and
// azimuthal differences drive the source terms, and near a body the
// potential's deviation from any ring profile is large, so an FP32 deviation
// would round the gradient there by about 2^-24 |dev| / |difference|.
__device__ __forceinline__ Fld<false> phi_fld(const KParams &P)
{
    return Fld<false>{static_cast<double *>(P.f[FCPT_POTENTIAL]), nullptr, P.ns, 0u};
}

// Neighbour values within a wavefront: the value of lane - 1 and of lane + 1,
// moved by the DPP wavefront shifts (wave_shr:1, wave_shl:1) on both 32-bit
// halves; the edge lanes, which have no source lane, receive 0 (bound_ctrl,
// so no old value has to be set first) and their results are discarded
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)(b & 0xffffffffll), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double from_left(double x) { return dpp_move<0x138>(x); }
__device__ __forceinline__ double from_right(double x) { return dpp_move<0x130>(x); }

// Kernels read every input before their first store: the compiler must
// assume that a store may alias any later load, so a load written after a
// store waits for it and the memory latency is paid once per such phase.

// Cell kernels run 64 x 4 blocks: one wavefront is one 64-sector piece of
// one ring, so the ring index is the same in every lane.  Reading it through
// readfirstlane tells the compiler so, and the per-ring metric values and
// backgrounds become scalar loads.
__device__ __forceinline__ int wave_ring(int r0, int by)
{
    return __builtin_amdgcn_readfirstlane(r0 + by * (int)blockDim.y + (int)threadIdx.y);
}

// XCD-aware block order.  The dispatcher hands workgroup b of a launch to XCD
// b mod 8 (each XCD has its own L2), so blocks adjacent in the grid, which
// share stencil halos (sector halo lanes, cache lines at segment edges, the
// rings a walk reloads), would run on different XCDs and fetch the shared
// lines once per XCD.  xcd_block gives each XCD a contiguous run of the grid
// instead: the workgroups XCD x receives (b = x, x + 8, ...) take the x-th
// run of linear indices in order, so neighbours share an L2 and run close in
// time.  A bijection of [0, gx gy); the block's work is unchanged.
struct BlockXY {
    int x, y;
};
__device__ __forceinline__ BlockXY xcd_block()
{
    constexpr unsigned kXcds = 8;
    const unsigned gx = gridDim.x, G = gridDim.x * gridDim.y;
    const unsigned b = blockIdx.y * gx + blockIdx.x;
    const unsigned xcd = b % kXcds, idx = b / kXcds, q = G / kXcds, r = G % kXcds;
    const unsigned L = xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
    return BlockXY{(int)(L % gx), (int)(L / gx)};
}

// 2-D launch over rows [r0, r1) x all sectors; (bx, by) is the block's place
// in the disk's own grid
#define CELL_LOOP_HEAD(r0, r1)                                   \
    const int ns = P.ns;                                         \
    const int j = bx * (int)blockDim.x + (int)threadIdx.x;       \
    const int i = wave_ring(r0, by);                             \
    if (j >= ns || i >= (r1)) return;

// Every kernel body takes the disk's parameters and its block coordinates
// (bx, by) in that disk's grid, and comes in two launch forms: k_NAME runs
// one disk on its own grid; kb_NAME runs many disks in one launch, each
// block reading (disk, by << 16 | bx) from a per-stage work table and the
// disk's parameters from a device array (constant address space: scalar
// loads).
__device__ __forceinline__ BatchEntry batch_entry(const BatchEntry *map)
{
    const auto *m = cst(map) + blockIdx.x;
    return BatchEntry{m->disk, m->yx};
}
__device__ __forceinline__ KParams batch_disk(const KParams *Ps, unsigned d)
{
    // a copy through the constant address space, word by word: the fields a
    // kernel uses become scalar loads at its start, as kernel arguments do
    static_assert(sizeof(KParams) % 8 == 0, "KParams is copied in 8-byte words");
    KParams P;
    const auto src = cst(reinterpret_cast<const unsigned long long *>(Ps + d));
    unsigned long long *dst = reinterpret_cast<unsigned long long *>(&P);
#pragma unroll
    for (unsigned k = 0; k < sizeof(KParams) / 8; ++k) dst[k] = src[k];
    return P;
}

#define FCPT_PAIR0(name, BOUNDS)                                                                  \
    template <bool SPLIT>                                                                         \
    __global__ void __launch_bounds__(BOUNDS) k_##name(KParams P)                                 \
    {                                                                                             \
        b_##name<SPLIT>(P, (int)blockIdx.x, (int)blockIdx.y);                                     \
    }                                                                                             \
    template <bool SPLIT>                                                                         \
    __global__ void __launch_bounds__(BOUNDS) kb_##name(const KParams *Ps, const BatchEntry *map)  \
    {                                                                                             \
        const BatchEntry e = batch_entry(map);                                                \
        b_##name<SPLIT>(batch_disk(Ps, e.disk), (int)(e.yx & 0xffffu), (int)(e.yx >> 16));        \
    }

#define FCPT_PAIR1(name, BOUNDS, T1, a1)                                                          \
    template <bool SPLIT>                                                                         \
    __global__ void __launch_bounds__(BOUNDS) k_##name(KParams P, T1 a1)                          \
    {                                                                                             \
        b_##name<SPLIT>(P, (int)blockIdx.x, (int)blockIdx.y, a1);                                 \
    }                                                                                             \
    template <bool SPLIT>                                                                         \
    __global__ void __launch_bounds__(BOUNDS)                                                     \
        kb_##name(const KParams *Ps, const BatchEntry *map, T1 a1)                                \
    {                                                                                             \
        const BatchEntry e = batch_entry(map);                                                \
        b_##name<SPLIT>(batch_disk(Ps, e.disk), (int)(e.yx & 0xffffu), (int)(e.yx >> 16), a1);    \
    }

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

// The ring sum and the sum of the magnitudes of its terms in one tree: the
// sum's additions are exactly those of block_sum<BS>.
template <int BS>
__device__ double2 block_sum_abs(double v, double a, double2 *sh)
{
    const int t = threadIdx.x;
    sh[t] = double2{v, a};
    __syncthreads();
#pragma unroll
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] = double2{sh[t].x + sh[t + s].x, sh[t].y + sh[t + s].y};
        __syncthreads();
    }
    return sh[0];
}

// FARGO shift of ring i from a ring sum (TransportEuler.cpp:215-220)
__device__ __forceinline__ double fargo_ntilde(const KParams &P, int i, double sum, double dt)
{
    const double vm = sum / (double)P.ns;
    return vm * cst(P.InvRmed)[i] * dt * P.invdphi;
}

// mode 0: CFL ring mean
// mode 1: transport ring mean + FARGO shift + uniform residual
//         (TransportEuler.cpp:207-236)
// The reference sums a ring left to right (TransportEuler.cpp:178-183); the
// block's tree sums the same terms in another order.  Any two summation
// orders of n terms differ by at most 2 gamma_(n-1) sum |term| (gamma_k =
// k u / (1 - k u), u = 2^-53), and the shift floor(Ntilde + 0.5) is a
// monotone function of the sum (a division by ns and products with positive
// ring constants, each rounded monotonically).  So when both ends of the
// interval sum -/+ that bound give the same shift, the left-to-right sum
// gives it too; otherwise thread 0 forms the left-to-right sum itself
// (seq_sum: the ring's terms in order) and the ring's mean, shift and
// residual come from it, as in the reference.
template <typename SeqSum>
__device__ __forceinline__ void ring_mean_store(const KParams &P, int i, double sum, double asum, int mode,
                                                SeqSum seq_sum)
{
    if (threadIdx.x != 0) return;
    if (mode == 1) {
        const double dt = P.s->dt;
        const double nsd = (double)P.ns;
        const double bound = asum * (2.02 * nsd * 0x1p-53) + fabs(sum) * 0x1p-50;
        const double lo = floor(fargo_ntilde(P, i, sum - bound, dt) + 0.5);
        const double hi = floor(fargo_ntilde(P, i, sum + bound, dt) + 0.5);
        if (!(lo == hi)) sum = seq_sum();
        const double vm = sum / nsd;
        P.vmean[i] = vm;
        const double invdt = 1.0 / dt;
        const double Ntilde = vm * cst(P.InvRmed)[i] * dt * P.invdphi;
        const double Nround = floor(Ntilde + 0.5);
        P.nshift[i] = (int)Nround;
        P.vuni[i] = (Ntilde - Nround) * cst(P.Rmed)[i] * invdt * P.dphi;
    } else {
        P.vmean[i] = sum / (double)P.ns;
    }
}

// One thread's share of a ring sum: the sectors t, t + STRIDE, t + 2 STRIDE,
// ... below ns, added in that order.  Eight sectors' loads are issued before
// their additions (a loop of one load and one addition keeps one load in
// flight per thread and waits for each): the same additions in the same
// order, so the same bits.
template <int STRIDE, class Load, class Add>
__device__ __forceinline__ void strided_row(int ns, int t, Load &&load, Add &&add)
{
    constexpr int U = 8;
    int j = t;
    for (; j + STRIDE * (U - 1) < ns; j += STRIDE * U) {
        double x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = load(j + STRIDE * u);
#pragma unroll
        for (int u = 0; u < U; ++u) add(x[u]);
    }
    for (; j < ns; j += STRIDE) add(load(j));
}

template <bool SPLIT>
__device__ __forceinline__ void b_ring_mean(const KParams &P, int bx, int by, int mode)
{
    __shared__ double2 sh[256];
    const int ns = P.ns;
    const int i = bx;
    FLD(v, FCPT_VAZI);
    double acc = 0.0, aacc = 0.0;
    strided_row<256>(ns, (int)threadIdx.x, [&](int j) { return v(i, j); }, [&](double x) {
        acc += x;
        aacc += fabs(x);
    });
    const double2 r = block_sum_abs<256>(acc, aacc, sh);
    ring_mean_store(P, i, r.x, r.y, mode, [&]() {
        double seq = 0.0;
        for (int j = 0; j < ns; ++j) seq += v(i, j);
        return seq;
    });
}
FCPT_PAIR1(ring_mean, 256, int, mode)

// cfl.cpp:210-376, per cell dt and the per-ring shear limit.  Each thread
// walks rows with a grid stride; the block minimum (wavefront shuffles, then
// LDS) goes to its own slot of P.ring_min and k_cfl_finalize folds the slots,
// so the result is deterministic and free of atomics.
// mode 0: condition_cfl only; 1: sim::CalculateTimeStep (simulation.cpp:100-118);
// 2: CalculateTimeStep + the monitor clip of sim::run (simulation.cpp:526-542).
// One block of 256 threads (thread index t) folds the block minima in a fixed
// pairwise order.
// With `rings` the slots are the per-ring minima of k_cfl_rings (nr of them)
// and the fold adds the per-ring shear limits from the ring means.
__device__ __forceinline__ double cfl_shear(const KParams &P, int k)
{
    // the limit of the shear between rings k and k + 1 (cfl.cpp:223-233)
    const double *vm = P.vmean;
    const double denom = fabs(vm[k] * P.InvRmed[k] - vm[k + 1] * P.InvRmed[k + 1]) + 1.0e-100;
    return P.cfl * P.dphi / denom;
}

__device__ __forceinline__ void b_cfl_fold(const KParams &P, int mode, int t, bool rings = false)
{
    const int nblocks = rings ? P.nr : P.cfl_gx * P.cfl_gy;
    __shared__ double sh[256];
    double m = __longlong_as_double(0x7FF0000000000000ll);
    for (int k = t; k < nblocks; k += 256) m = std_min(m, P.ring_min[k]);
    if (rings)
        for (int k = t; k < P.nr - 1; k += 256) m = std_min(m, cfl_shear(P, k));
    sh[t] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) sh[t] = std_min(sh[t], sh[t + w]);
        __syncthreads();
    }
    if (t != 0) return;
    DevScalars *s = P.s;
    const double cfl_dt = sh[0];
    s->cfl_dt = cfl_dt;
    s->step_count += 1u;
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
__device__ __forceinline__ void b_cfl_finalize(const KParams &P, int mode) { b_cfl_fold(P, mode, (int)threadIdx.x); }

// adiabatic sound speed, scale height and viscosity of one cell
// (compute_sound_speed, compute_scale_height, compute_viscosity)
struct AdiDerived {
    double cs, h, nu;
};
template <bool SPLIT>
__device__ __forceinline__ AdiDerived adi_derived(const KParams &P, int i, double e, double sig)
{
    AdiDerived d;
    const double g = P.gamma;
    const double inv_omega_kepler = ring_tab(P, TAB_INV_OMEGA_K, i);  // 1 / Omega_K(Rmed)
    if constexpr (SPLIT) {  // products with reciprocals (the 1 / sqrt(gamma) is launch-wide)
        d.cs = sqrt(g * (g - 1.0) * e * recip<true>(sig));
        d.h = d.cs * inv_sqrt<true>(g) * inv_omega_kepler;
    } else {
        d.cs = sqrt_ieee(g * (g - 1.0) * e / sig);
        d.h = d.cs / (sqrt(g)) * inv_omega_kepler;
    }
    const double alpha = P.alpha;
    d.nu = alpha * d.h * d.cs;
    return d;
}

// true where the fused step forms c_s, nu and T instead of storing them
template <bool SPLIT>
__device__ __forceinline__ bool lazy_derived(const KParams &P)
{
    return P.lazy_derived != 0 && P.adiabatic;
}

// the CFL limit of cell (i, j) (cfl.cpp:235-376), vm_i the ring's mean v_phi
template <bool SPLIT>
__device__ __forceinline__ double cfl_cell(const KParams &P, int i, int j, double vm_i)
{
    const int ns = P.ns;
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(cs, FCPT_SOUNDSPEED);
    FLD(nu, FCPT_VISCOSITY);
    const bool iso = !P.adiabatic;  // c_s and nu are ring profiles
    const double CFL = P.cfl, dphi = P.dphi;
    const double fac2 = P.artvisc_factor * P.artvisc_factor;
    const double dxRadial = cst(P.Rsup)[i] - cst(P.Rinf)[i];
    const double dxAzimuthal = cst(P.Rmed)[i] * dphi;
    const double cell_size = std_min(dxRadial, dxAzimuthal);
    const int jn = j == ns - 1 ? 0 : j + 1;
    const double vr0 = vr(i, j), vr1 = vr(i + 1, j), va0 = vaz(i, j), va1 = vaz(i, jn);
    const double v_residual = va0 - vm_i;
    // split storage: divisions by the ring's cell sizes become products
    const double inv_dr = SPLIT ? cst(P.InvDiffRsup)[i] : 0.0;
    const double inv_daz = SPLIT ? recip<true>(dxAzimuthal) : 0.0;
    const double inv_cell = SPLIT ? recip<true>(cell_size) : 0.0;
    double csv, nuv;
    if (iso) {
        csv = cst(P.iso_cs)[i];
        nuv = cst(P.iso_nu)[i];
    } else if (lazy_derived<SPLIT>(P)) {
        // as the end of the last step formed them from the same E and Sigma
        const AdiDerived a = adi_derived<SPLIT>(P, i, fld<SPLIT>(P, FCPT_ENERGY)(i, j), fld<SPLIT>(P, FCPT_SIGMA)(i, j));
        csv = a.cs;
        nuv = a.nu;
    } else {
        csv = cs(i, j);
        nuv = nu(i, j);
    }
    const double invdt1 = SPLIT ? csv * inv_cell : csv / cell_size;
    const double invdt2 = SPLIT ? vr0 * inv_dr : vr0 / dxRadial;
    const double invdt3 = SPLIT ? v_residual * inv_daz : v_residual / dxAzimuthal;
    // the LeapFrog scheme applies the source terms twice with dt / 2 and
    // relaxes their limits (cfl.cpp:254-261); the Euler factor 1.0 is exact
    const double lf = P.integrator == FCPT_INTEGRATOR_LEAPFROG ? 0.6 : 1.0;
    double invdt4;
    if (P.artvisc == FCPT_ARTVISC_SN) {
        double dvRadial = vr1 - vr0;
        double dvAzimuthal = va1 - va0;
        if (dvRadial > 0.0) dvRadial = 0.0;
        else dvRadial = -dvRadial;
        if (dvAzimuthal > 0.0) dvAzimuthal = 0.0;
        else dvAzimuthal = -dvAzimuthal;
        if constexpr (SPLIT)
            invdt4 = 4.0 * fac2 * std_max(dvRadial * inv_dr, dvAzimuthal * inv_daz) * lf;
        else
            invdt4 = 4.0 * fac2 * std_max(dvRadial / dxRadial, dvAzimuthal / dxAzimuthal) * lf;
    } else {
        const double eps_rr = (vr1 - vr0) * cst(P.InvDiffRsup)[i];
        const double eps_pp = cst(P.InvRmed)[i] * ((va1 - va0) * P.invdphi + 0.5 * (vr1 + vr0));
        const double mdiv_V = -std_min(eps_rr + eps_pp, 0.0);
        invdt4 = 4.0 * fac2 * mdiv_V * lf;
    }
    const double invdt5 = SPLIT ? 4.0 * nuv * (inv_cell * inv_cell) * lf
                                : 4.0 * nuv / (cell_size * cell_size) * lf;
    double invdt6 = 0.0;
    if (P.adiabatic) {
        FLD(qp, FCPT_QPLUS);
        FLD(qm, FCPT_QMINUS);
        FLD(E, FCPT_ENERGY);
        const double inv_limit = 1.0 / P.heating_cooling_cfl_limit;
        if constexpr (SPLIT) invdt6 = inv_limit * fabs((qp(i, j) - qm(i, j)) * recip<true>(E(i, j))) * lf;
        else invdt6 = inv_limit * fabs((qp(i, j) - qm(i, j)) / E(i, j)) * lf;
    }
    const double sum2 = invdt1 * invdt1 + invdt2 * invdt2 + invdt3 * invdt3 + invdt4 * invdt4 +
                        invdt5 * invdt5 + invdt6 * invdt6;
    const double dt_cell = SPLIT ? CFL * inv_sqrt<true>(sum2) : CFL / sqrt(sum2);
    return dt_cell;
}

// per-cell CFL limits (cfl.cpp:185-382) and this block's minimum in its slot
// (k_cfl_finalize folds the slots: a device-wide fold inside this kernel
// would need a device-scope release per block, which on gfx950 writes back
// the XCD's L2 and costs more than the separate launch)
template <bool SPLIT>
__device__ __forceinline__ void b_cfl_cells(const KParams &P, int bx, int by)
{
    const int ns = P.ns;
    const int j = bx * blockDim.x + threadIdx.x;
    double cand = __longlong_as_double(0x7FF0000000000000ll);
    const auto vm = cst(P.vmean);
    for (int i = wave_ring(1, by); i < P.nr - 1 && j < ns; i += P.cfl_gy * (int)blockDim.y) {
        if (j == 0) {
            cand = std_min(cand, cfl_shear(P, i));
            if (i == 1) cand = std_min(cand, cfl_shear(P, 0));
        }
        // NaN never wins a comparison, exactly like the reference's "if (dt_cell < dt_core)"
        cand = std_min(cand, cfl_cell<SPLIT>(P, i, j, vm[i]));
    }
    for (int off = 32; off > 0; off >>= 1) cand = std_min(cand, __shfl_xor(cand, off, 64));
    __shared__ double wmin[4];
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    if ((t & 63) == 0) wmin[t >> 6] = cand;
    __syncthreads();
    if (t == 0) {
        const double m = std_min(std_min(wmin[0], wmin[1]), std_min(wmin[2], wmin[3]));
        P.ring_min[by * P.cfl_gx + bx] = m;
    }
}
FCPT_PAIR0(cfl_cells, 256)

// The CFL ring mean and cell limits in one launch, one block of 256 threads
// per ring: the block forms the ring's mean v_phi exactly as k_ring_mean does
// (same per-thread sums, same tree) and stores it, then the limits of the
// ring's cells (rings 1 .. nr-2) with that mean, and puts the block minimum in
// the ring's slot; the fold (rings = true) adds the shear limits between
// neighbouring rings' means.  The minimum over the same values as
// k_ring_mean + k_cfl_cells form.
// 512 threads: the first 256 form the mean (the 256-thread sums and tree of
// k_ring_mean; the others take part in the barriers only), all 512 the cells.
template <bool SPLIT>
__device__ __forceinline__ void b_cfl_rings(const KParams &P, int i)
{
    __shared__ double sh[256];
    __shared__ double wmin[8];
    const int ns = P.ns;
    const int t = (int)threadIdx.x;
    FLD(v, FCPT_VAZI);
    double acc = 0.0;
    if (t < 256) strided_row<256>(ns, t, [&](int j) { return v(i, j); }, [&](double x) { acc += x; });
    if (t < 256) sh[t] = acc;
    __syncthreads();
#pragma unroll
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) sh[t] = sh[t] + sh[t + w];
        __syncthreads();
    }
    const double vm = sh[0] / (double)ns;
    if (t == 0) P.vmean[i] = vm;
    double cand = __longlong_as_double(0x7FF0000000000000ll);
    if (i >= 1 && i < P.nr - 1)
        for (int j = t; j < ns; j += 512) cand = std_min(cand, cfl_cell<SPLIT>(P, i, j, vm));
    for (int off = 32; off > 0; off >>= 1) cand = std_min(cand, __shfl_xor(cand, off, 64));
    if ((t & 63) == 0) wmin[t >> 6] = cand;
    __syncthreads();
    if (t == 0) {
        double m = wmin[0];
        for (int k = 1; k < 8; ++k) m = std_min(m, wmin[k]);
        P.ring_min[i] = m;
    }
}
template <bool SPLIT>
__global__ void __launch_bounds__(512) k_cfl_rings(KParams P)
{
    b_cfl_rings<SPLIT>(P, (int)blockIdx.x);
}
template <bool SPLIT>
__global__ void __launch_bounds__(512) kb_cfl_rings(const KParams *Ps, const BatchEntry *map)
{
    const BatchEntry e = batch_entry(map);
    b_cfl_rings<SPLIT>(batch_disk(Ps, e.disk), (int)(e.yx & 0xffffu));
}

__global__ void __launch_bounds__(256) k_cfl_finalize(KParams P, int mode) { b_cfl_finalize(P, mode); }
__global__ void __launch_bounds__(256) k_cfl_finalize_rings(KParams P, int mode)
{
    b_cfl_fold(P, mode, (int)threadIdx.x, true);
}
__global__ void __launch_bounds__(256) kb_cfl_finalize_rings(const KParams *Ps, const BatchEntry *map, int mode)
{
    b_cfl_fold(batch_disk(Ps, batch_entry(map).disk), mode, (int)threadIdx.x, true);
}
__global__ void __launch_bounds__(256) kb_cfl_finalize(const KParams *Ps, const BatchEntry *map, int mode)
{
    b_cfl_finalize(batch_disk(Ps, batch_entry(map).disk), mode);
}

__global__ void k_set_dt(KParams P, double dt) { P.s->dt = dt; }
__global__ void k_set_clock(KParams P, double time) { P.s->time = time; }

__global__ void k_set_time(KParams P, double time, double last_dt, int n_monitor)
{
    P.s->time = time;
    P.s->last_dt = last_dt;
    P.s->n_monitor = n_monitor;
}

__global__ void k_set_params(KParams P, fcpt_step_params sp, BodySmoothing sm)
{
    DevScalars *s = P.s;
    s->n_planets = sp.n_planets;
    for (int k = 0; k < sp.n_planets; ++k) {
        s->planet[k].x = sp.planet[k].x;
        s->planet[k].y = sp.planet[k].y;
        s->planet[k].mass = sp.planet[k].mass;
        s->planet[k].r_cubic = sp.planet[k].r_cubic;
        s->planet[k].smooth = sm.s[k];
    }
    s->indirect_x = sp.indirect_x;
    s->indirect_y = sp.indirect_y;
    s->has_mid = 0;
}

// the bodies of a LeapFrog step's second kick
__global__ void k_set_mid_params(KParams P, fcpt_step_params sp, BodySmoothing sm)
{
    DevScalars *s = P.s;
    s->n_planets_mid = sp.n_planets;
    for (int k = 0; k < sp.n_planets; ++k) {
        s->planet_mid[k].x = sp.planet[k].x;
        s->planet_mid[k].y = sp.planet[k].y;
        s->planet_mid[k].mass = sp.planet[k].mass;
        s->planet_mid[k].r_cubic = sp.planet[k].r_cubic;
        s->planet_mid[k].smooth = sm.s[k];
    }
    s->indirect_mid_x = sp.indirect_x;
    s->indirect_mid_y = sp.indirect_y;
    s->has_mid = 1;
}

// step_LeapFrog's clock (simulation.cpp:278-281, 378, 441): phase 0 before
// the first kick (dt / 2 at the start time), 1 before the drift (dt), 2 before
// the second kick (dt / 2 at the midstep time, the midstep bodies), 3 before
// the final boundaries (dt at the start time: they advance the time to
// start + dt, the reference's end_time)
__global__ void k_leapfrog_phase(KParams P, int phase)
{
    DevScalars *s = P.s;
    if (phase == 0) {
        s->dt_step = s->dt;
        s->t_step = s->time;
        s->dt = s->dt_step / 2;
    } else if (phase == 1) {
        s->dt = s->dt_step;
    } else if (phase == 2) {
        s->dt = s->dt_step / 2;
        s->time = s->t_step + s->dt;
        if (s->has_mid) {
            s->n_planets_first = s->n_planets;
            s->indirect_first_x = s->indirect_x;
            s->indirect_first_y = s->indirect_y;
            for (int k = 0; k < kMaxPlanets; ++k) s->planet_first[k] = s->planet[k];
            s->n_planets = s->n_planets_mid;
            s->indirect_x = s->indirect_mid_x;
            s->indirect_y = s->indirect_mid_y;
            for (int k = 0; k < kMaxPlanets; ++k) s->planet[k] = s->planet_mid[k];
        }
    } else {
        s->dt = s->dt_step;
        s->time = s->t_step;
        if (s->has_mid) {
            s->n_planets = s->n_planets_first;
            s->indirect_x = s->indirect_first_x;
            s->indirect_y = s->indirect_first_y;
            for (int k = 0; k < kMaxPlanets; ++k) s->planet[k] = s->planet_first[k];
        }
    }
}

// ---------------------------------------------------------------------------
// x^4 and x^3 of the Klahr-Kley factor by compensated powering (Graillat, Langlois and Louvet,
// "Algorithms for accurate, validated and fast polynomial evaluation",
// Japan J. Indust. Appl. Math. 26, 2009): the square is carried as the
// unevaluated sum p + e (e from an FMA), and the last product's rounding error
// is added back before the one final rounding, so the result is within 0.5 +
// 2^-49 ulp of the exact power: the correctly rounded value but for inputs
// within 2^-49 ulp of a rounding boundary, as the C library's pow (within
// 0.52 ulp) returns it.  An overflowing power returns the rounded product.
__device__ __forceinline__ double pow4_comp(double x)
{
    const double p = x * x, e = fma(x, x, -p);
    const double q = p * p, eq = fma(p, p, -q);
    const double r = q + (eq + 2.0 * p * e);
    return isfinite(q) ? r : q;
}
__device__ __forceinline__ double pow3_comp(double x)
{
    const double p = x * x, e = fma(x, x, -p);
    const double q = p * x, eq = fma(p, x, -q);
    const double r = q + (eq + e * x);
    return isfinite(q) ? r : q;
}

// one body's term at (x, y) with smoothing length `smooth`
template <bool SPLIT>
__device__ __forceinline__ double potential_body(const KParams &P, double x, double y, double smooth,
                                                 double px, double py, double mass, double r_sm)
{
    const double dx = x - px;
    const double dy = y - py;
    const double dist_2 = dx * dx + dy * dy;
    const double d2 = dist_2 + smooth * smooth;
    // split storage: one reciprocal square root instead of sqrt and division
    const double inv_d = SPLIT ? inv_sqrt<true>(d2) : 0.0;
    const double d_smoothed = SPLIT ? d2 * inv_d : sqrt_ieee(d2);
    double smooth_factor_klahr = 1.0;
    if (r_sm > 0.0 && d_smoothed < r_sm) {
        const double q = d_smoothed / r_sm;
        smooth_factor_klahr = (pow4_comp(q) - 2.0 * pow3_comp(q) + 2.0 * d_smoothed / r_sm);
    }
    if constexpr (SPLIT) return -P.G * mass * inv_d * smooth_factor_klahr;
    else return shared_div(-P.G * mass, shared_den(d_smoothed)) * smooth_factor_klahr;
}

// CalculateNbodyPotential, Pframeforce.cpp:44-85 (+ compute_smoothing,
// Force.cpp:145-159: a body's fixed smoothing length, or ThicknessSmoothing *
// H of the cell), for kPotCells cells of ring i, sectors j0 + 64 c (those
// below ns), with the bodies' loop outside the cells': every cell still adds the bodies
// in order and then the indirect term, and a body's parameters are loaded
// once per kPotCells cells.  The potential is stored in FP64 in both storage
// modes.
constexpr int kPotCells = 4;
template <bool SPLIT>
__device__ __forceinline__ void potential_ring_cells(const KParams &P, int i, int j0)
{
    const DevScalars *s = P.s;
    const int ns = P.ns;
    double x[kPotCells], y[kPotCells], sm[kPotCells], acc[kPotCells];
#pragma unroll
    for (int c = 0; c < kPotCells; ++c) {
        const int j = min(j0 + 64 * c, ns - 1);  // past ns: formed, not stored
        x[c] = cst(P.Rmed)[i] * P.cosphi[j];
        y[c] = cst(P.Rmed)[i] * P.sinphi[j];
        const double H = P.adiabatic ? fld<SPLIT>(P, FCPT_SCALE_HEIGHT)(i, j) : cst(P.iso_h)[i];
        sm[c] = P.thickness_smoothing * H;
        acc[c] = 0.0;
    }
    // the bodies through scalar loads (this launch writes none of them)
    const auto bodies = cst(s->planet);
    const int np = *cst(&s->n_planets);
    for (int k = 0; k < np; ++k) {
        const double px = bodies[k].x, py = bodies[k].y;
        const double mass = bodies[k].mass, r_sm = bodies[k].r_cubic;
        const double smk = bodies[k].smooth;
#pragma unroll
        for (int c = 0; c < kPotCells; ++c)
            acc[c] += potential_body<SPLIT>(P, x[c], y[c], smk >= 0.0 ? smk : sm[c], px, py, mass, r_sm);
    }
    const double ix = *cst(&s->indirect_x), iy = *cst(&s->indirect_y);
#pragma unroll
    for (int c = 0; c < kPotCells; ++c) {
        const int j = j0 + 64 * c;
        if (j >= ns) break;
        phi_fld(P).set(i, j, acc[c] + (-ix * x[c] - iy * y[c]));
    }
}

// the potential, blocks of 64 x kPotCells sectors by 4 rings
template <bool SPLIT>
__device__ __forceinline__ void b_potential(const KParams &P, int bx, int by)
{
    const int i = wave_ring(0, by);
    const int j0 = bx * 64 * kPotCells + (int)threadIdx.x;
    if (j0 >= P.ns || i >= P.nr) return;
    potential_ring_cells<SPLIT>(P, i, j0);
}
FCPT_PAIR0(potential, 256)

// The potential of a single disk in one launch (b_potential); fold_mode >= 0:
// one extra block row runs the CFL fold of the step's start
// (k_cfl_finalize_rings; the potential reads none of the scalars the fold
// writes, and dt is first read by the kernels after this one).
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_potential_fold(KParams P, int fold_mode)
{
    // block row 0 is the fold (dispatched first, it ends under the others)
    const int bx = (int)blockIdx.x, by = (int)blockIdx.y - 1;
    if (by < 0) {
        if (bx == 0 && fold_mode >= 0)
            b_cfl_fold(P, fold_mode, (int)(threadIdx.y * blockDim.x + threadIdx.x), true);
        return;
    }
    b_potential<SPLIT>(P, bx, by);
}

// ---------------------------------------------------------------------------
// momentum_update_radial, SourceEuler.cpp:325-372, for one cell: pressure and
// potential gradients between rings i - 1 and i, and the centrifugal term of
// the four v_phi around the interface (summed in that order)
template <bool SPLIT = false>
__device__ __forceinline__ double src_vr_new(const KParams &P, int i, double dt, double v, double s0,
                                             double sm, double p0, double pm, double phi0, double phim,
                                             double va00, double va01, double va10, double va11)
{
    double gradp = quot<SPLIT>(2.0, s0 + sm);
    gradp *= (p0 - pm);
    gradp *= cst(P.InvDiffRmed)[i];
    const double gradphi = (phi0 - phim) * cst(P.InvDiffRmed)[i];
    const double vsum = va00 + va01 + va10 + va11;
    const double vt = 0.25 * vsum + cst(P.Rinf)[i] * P.omega_frame;
    const double vt2 = vt * vt;
    const double centrifugal_accel = vt2 * cst(P.InvRinf)[i];
    return v + dt * (-gradp - gradphi + centrifugal_accel);
}

// momentum_update_azimuthal, SourceEuler.cpp:375-428, for one cell
template <bool SPLIT = false>
__device__ __forceinline__ double src_vphi_new(const KParams &P, int i, double dt, double va, double s0,
                                               double sjp, double p0, double pjp, double phi0,
                                               double phijp)
{
    const double invdxtheta = ring_tab(P, TAB_SRC_INVDXTHETA, i);  // 2 / (dphi (Rsup + Rinf))
    const double gradp = quot<SPLIT>(2.0, s0 + sjp) * (p0 - pjp) * invdxtheta;
    const double gradphi = (phi0 - phijp) * invdxtheta;
    return va + dt * (-gradp - gradphi);
}

// momentum_update_radial, SourceEuler.cpp:325-372
template <bool SPLIT>
__device__ __forceinline__ void b_src_vr(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(2, P.nr - 1)
    FLD(S, FCPT_SIGMA);
    FLD(Pr, FCPT_PRESSURE);
    const Fld<false> Phi = phi_fld(P);
    FLD(vaz, FCPT_VAZI);
    FLD(vr, FCPT_VRAD);
    const double dt = P.s->dt;
    const int jn = j == ns - 1 ? 0 : j + 1;
    vr.set(i, j, src_vr_new<SPLIT>(P, i, dt, vr(i, j), S(i, j), S(i - 1, j), Pr(i, j), Pr(i - 1, j), Phi(i, j),
                            Phi(i - 1, j), vaz(i, j), vaz(i, jn), vaz(i - 1, j), vaz(i - 1, jn)));
}
FCPT_PAIR0(src_vr, 256)

// momentum_update_azimuthal, SourceEuler.cpp:375-428
template <bool SPLIT>
__device__ __forceinline__ void b_src_vphi(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(1, P.nr - 1)
    FLD(S, FCPT_SIGMA);
    FLD(Pr, FCPT_PRESSURE);
    const Fld<false> Phi = phi_fld(P);
    FLD(vaz, FCPT_VAZI);
    const double dt = P.s->dt;
    const int jp = j == 0 ? ns - 1 : j - 1;
    vaz.set(i, j, src_vphi_new<SPLIT>(P, i, dt, vaz(i, j), S(i, j), S(i, jp), Pr(i, j), Pr(i, jp), Phi(i, j),
                               Phi(i, jp)));
}
FCPT_PAIR0(src_vphi, 256)

// compression_heating, SourceEuler.cpp:459-493, for one cell (rings 0 .. nr-2)
// from v_r(i), v_r(i + 1), v_phi(j), v_phi(j + 1)
__device__ __forceinline__ double src_energy_new(const KParams &P, int i, double dt, double e, double vr0,
                                                 double vr1, double va0, double va1)
{
    const double DIV_V = (vr1 * cst(P.Rinf)[i + 1] - vr0 * cst(P.Rinf)[i]) * cst(P.InvDiffRsupRb)[i] +
                         (va1 - va0) * P.invdphi * cst(P.InvRmed)[i];
    return e * exp(-(P.gamma - 1.0) * dt * DIV_V);
}

template <bool SPLIT>
__device__ __forceinline__ void b_src_energy(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr - 1)
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(E, FCPT_ENERGY);
    const double dt = P.s->dt;
    const int jn = j == ns - 1 ? 0 : j + 1;
    E.set(i, j, src_energy_new(P, i, dt, E(i, j), vr(i, j), vr(i + 1, j), vaz(i, j), vaz(i, jn)));
}
FCPT_PAIR0(src_energy, 256)

// the energy of temperature t, t sigma / mu R / (gamma - 1) (split storage:
// one product with the launch-wide R / (mu (gamma - 1)))
template <bool SPLIT = false>
__device__ __forceinline__ double energy_at(const KParams &P, double t, double sigma)
{
    if constexpr (SPLIT) return t * sigma * (P.Rgas * recip<true>(P.mu * (P.gamma - 1.0)));
    else return t * sigma / P.mu * P.Rgas / (P.gamma - 1.0);
}
template <bool SPLIT = false>
__device__ __forceinline__ double energy_floor(const KParams &P, double sigma)
{
    return energy_at<SPLIT>(P, P.t_min, sigma);
}
// assure_temperature_range for one cell, SourceEuler.cpp:136-202
template <bool SPLIT = false>
__device__ __forceinline__ double temperature_floor(const KParams &P, double e, double sigma)
{
    const double minimum_energy = energy_floor<SPLIT>(P, sigma);
    const double maximum_energy = energy_at<SPLIT>(P, P.t_max, sigma);
    if (!(e > minimum_energy)) e = minimum_energy;
    if (!(e < maximum_energy)) e = maximum_energy;
    return e;
}

template <bool SPLIT>
__device__ __forceinline__ void b_temperature_floor(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(E, FCPT_ENERGY);
    FLD(S, FCPT_SIGMA);
    E.set(i, j, temperature_floor<SPLIT>(P, E(i, j), S(i, j)));
}
FCPT_PAIR0(temperature_floor, 256)

// ---------------------------------------------------------------------------
// compute_sound_speed / compute_scale_height / update_viscosity for one cell
// (SourceEuler.cpp:957-995, 1121-1154; viscosity.cpp:98-137), as a load of
// the inputs and a store of the results
// locally isothermal pressure Sigma c_s^2 (the value recalculate_derived
// stores, SourceEuler.cpp:225-249, with the ring profile of c_s)
__device__ __forceinline__ double iso_pressure(const KParams &P, int i, double sig)
{
    const double c_s = cst(P.iso_cs)[i];
    return sig * (c_s * c_s);
}

struct DerivedIn {
    double e, sig, cs, h;
};

template <bool SPLIT>
__device__ __forceinline__ DerivedIn derived_load(const KParams &P, int i, int j, bool with_pressure)
{
    DerivedIn d{0.0, 0.0, 0.0, 0.0};
    if (P.adiabatic) {
        d.e = fld<SPLIT>(P, FCPT_ENERGY)(i, j);
        d.sig = fld<SPLIT>(P, FCPT_SIGMA)(i, j);
    } else {
        // locally isothermal: c_s and H are the ring profiles of k_iso_rings
        d.cs = cst(P.iso_cs)[i];
        d.h = cst(P.iso_h)[i];
        if (with_pressure) d.sig = fld<SPLIT>(P, FCPT_SIGMA)(i, j);
    }
    return d;
}

template <bool SPLIT>
__device__ __forceinline__ void derived_store(const KParams &P, int i, int j, const DerivedIn &d,
                                             bool with_temperature, bool with_pressure,
                                             bool force_nu = false)
{
    double c_s, h;
    // the fused step's kick walk forms the pressure from E itself
    if (P.e_partner) with_pressure = false;
    if (lazy_derived<SPLIT>(P)) {
        // the field the step reads stored: H (the potential's smoothing)
        const AdiDerived a = adi_derived<SPLIT>(P, i, d.e, d.sig);
        fld<SPLIT>(P, FCPT_SCALE_HEIGHT).set(i, j, a.h);
        if (with_pressure) fld<SPLIT>(P, FCPT_PRESSURE).set(i, j, (P.gamma - 1.0) * d.e);
        return;
    }
    if (P.adiabatic) {
        const double e = d.e;
        const double sig = d.sig;
        if (with_temperature) {
            const double c_v_inv = P.mu / P.Rgas * (P.gamma - 1.0);
            fld<SPLIT>(P, FCPT_TEMPERATURE).set(i, j, c_v_inv * e / sig);
        }
        const AdiDerived a = adi_derived<SPLIT>(P, i, e, sig);
        c_s = a.cs;
        h = a.h;
        fld<SPLIT>(P, FCPT_SOUNDSPEED).set(i, j, c_s);
        fld<SPLIT>(P, FCPT_SCALE_HEIGHT).set(i, j, h);
        if (with_pressure) fld<SPLIT>(P, FCPT_PRESSURE).set(i, j, (P.gamma - 1.0) * e);
    } else {
        c_s = d.cs;
        h = d.h;
        if (with_pressure) fld<SPLIT>(P, FCPT_PRESSURE).set(i, j, d.sig * (c_s * c_s));
        // locally isothermal: nu = alpha H c_s is the ring profile iso_nu,
        // already in the field since the state was uploaded
        if (!force_nu) return;
    }
    const double alpha = P.alpha;
    fld<SPLIT>(P, FCPT_VISCOSITY).set(i, j, alpha * h * c_s);
}

// ---------------------------------------------------------------------------
// The tensor of the TW artificial viscosity at one cell from its velocities
// v_r(i), v_r(i + 1), v_phi(j), v_phi(j + 1) (artificial_viscosity.cpp:44-70)
struct TwQ {
    double qr, qp, eps_rr, eps_pp, div_V, l_sq;
};

__device__ __forceinline__ TwQ tw_q_cell(const KParams &P, int i, double vr0, double vr1, double va0,
                                         double va1, double sig)
{
    const int ns = P.ns;
    TwQ q;
    q.eps_rr = (vr1 - vr0) * cst(P.InvDiffRsup)[i];
    q.eps_pp = cst(P.InvRmed)[i] * ((va1 - va0) * P.invdphi + 0.5 * (vr1 + vr0));
    q.div_V = std_min(q.eps_rr + q.eps_pp, 0.0);
    const double Dr = cst(P.Rinf)[i + 1] - cst(P.Rinf)[i];
    const double rDphi = cst(P.Rmed)[i] * P.dphi;
    const double dxm = (ns <= 16) ? std_min(Dr, rDphi) : std_max(Dr, rDphi);
    const double dx_sq = dxm * dxm;
    q.l_sq = (P.artvisc_factor * P.artvisc_factor) * dx_sq;
    q.qr = q.l_sq * sig * -q.div_V * (q.eps_rr - 1.0 / 3.0 * q.div_V);
    q.qp = q.l_sq * sig * -q.div_V * (q.eps_pp - 1.0 / 3.0 * q.div_V);
    return q;
}

// energy after the TW dissipation and the temperature floor
// (artificial_viscosity.cpp:72-86, then SetTemperatureFloorCeilValues, :19-21)
__device__ __forceinline__ double tw_dissipation(const KParams &P, int i, const TwQ &q, double sig,
                                                 double dt, double e)
{
    if (i > 1 && i < P.nr - 1) {
        const double d = q.eps_rr - q.eps_pp;
        const double Qplus =
            -q.l_sq * q.div_V * sig * 1.0 / 3.0 * (q.eps_rr * q.eps_rr + q.eps_pp * q.eps_pp + d * d);
        e += Qplus * dt;
    }
    return temperature_floor(P, e, sig);
}

// update_with_artificial_viscosity_TW, artificial_viscosity.cpp:35-88
template <bool SPLIT>
__device__ __forceinline__ void b_tw_q(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(S, FCPT_SIGMA);
    FLD(E, FCPT_ENERGY);
    const bool diss = P.adiabatic && P.artvisc_dissipation;
    const double dt = P.s->dt;
    const int jn = j == ns - 1 ? 0 : j + 1;
    const double vr0 = vr(i, j), vr1 = vr(i + 1, j), va0 = vaz(i, j), va1 = vaz(i, jn);
    const double sig = S(i, j);
    const double e_in = diss ? E(i, j) : 0.0;
    const TwQ q = tw_q_cell(P, i, vr0, vr1, va0, va1, sig);
    fld<SPLIT>(P, FCPT_Q_R).set(i, j, q.qr);
    fld<SPLIT>(P, FCPT_Q_PHI).set(i, j, q.qp);
    if (diss) E.set(i, j, tw_dissipation(P, i, q, sig, dt, e_in));
}
FCPT_PAIR0(tw_q, 256)

// the TW velocity updates of one cell from the tensors around it
// (artificial_viscosity.cpp:90-139)
template <bool SPLIT = false>
__device__ __forceinline__ double tw_vphi_new(const KParams &P, int i, double dt, double va, double s0,
                                              double sjp, double qp0, double qpj)
{
    const double sigma_phi_avg = 0.5 * (s0 + sjp);
    const double dVp =
        quot<SPLIT>(2.0 * dt, (cst(P.Rsup)[i] + cst(P.Rinf)[i]) * sigma_phi_avg) * (qp0 - qpj) * P.invdphi;
    return va + dVp;
}

template <bool SPLIT = false>
__device__ __forceinline__ double tw_vr_new(const KParams &P, int i, double dt, double v, double s0,
                                            double sm, double qr0, double qrm, double qp0, double qpm)
{
    const double sigma_r_avg = 0.5 * (s0 + sm);
    const double r0 = cst(P.Rmed)[i], r1 = cst(P.Rmed)[i - 1];
    const double dVr = quot<SPLIT>(quot<SPLIT>(P.radial_viscosity_factor * dt, sigma_r_avg) * 2.0, r0 * r0 - r1 * r1) *
                       ((qr0 * r0 - qrm * r1) - 0.5 * (qp0 + qpm) * (r0 - r1));
    return v + dVr;
}

template <bool SPLIT>
__device__ __forceinline__ void b_tw_vel(const KParams &P, int bx, int by, int with_viscosity)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(S, FCPT_SIGMA);
    FLD(Qr, FCPT_Q_R);
    FLD(Qp, FCPT_Q_PHI);
    FLD(vaz, FCPT_VAZI);
    FLD(vr, FCPT_VRAD);
    const double dt = P.s->dt;
    const bool inner = i >= 1 && i < P.nr - 1;
    const bool radial = i >= 2 && inner;
    const int jp = j == 0 ? ns - 1 : j - 1;
    const int im = i >= 1 ? i - 1 : 0;
    // in the fused step, recalculate_viscosity (SourceEuler.cpp:205-223), which is
    // pointwise in Sigma and E and follows this kernel, is done here for every ring
    DerivedIn din{0.0, 0.0, 0.0, 0.0};
    const bool visc = with_viscosity && P.adiabatic;  // constant ring profile when isothermal
    if (visc) din = derived_load<SPLIT>(P, i, j, false);
    double s0 = 0.0, sjp = 0.0, sm = 0.0, qp0 = 0.0, qpj = 0.0, qpm = 0.0, qr0 = 0.0, qrm = 0.0;
    double va = 0.0, v = 0.0;
    if (inner) {
        s0 = S(i, j);
        sjp = S(i, jp);
        qp0 = Qp(i, j);
        qpj = Qp(i, jp);
        va = vaz(i, j);
    }
    if (radial) {
        sm = S(im, j);
        qr0 = Qr(i, j);
        qrm = Qr(im, j);
        qpm = Qp(im, j);
        v = vr(i, j);
    }
    if (visc) derived_store<SPLIT>(P, i, j, din, false, false);
    if (!inner) return;
    vaz.set(i, j, tw_vphi_new<SPLIT>(P, i, dt, va, s0, sjp, qp0, qpj));
    if (radial) vr.set(i, j, tw_vr_new<SPLIT>(P, i, dt, v, s0, sm, qr0, qrm, qp0, qpm));
}
FCPT_PAIR1(tw_vel, 256, int, with_viscosity)

// The von Neumann-Richtmyer tensor at one cell (artificial_viscosity.cpp:160-185)
struct SnQ {
    double qr, qp, dv_r, dv_phi;
};

__device__ __forceinline__ SnQ sn_q_cell(const KParams &P, double vr0, double vr1, double va0,
                                         double va1, double sig)
{
    SnQ q;
    const double fac2 = P.artvisc_factor * P.artvisc_factor;
    q.dv_r = vr1 - vr0;
    q.qr = q.dv_r < 0.0 ? fac2 * sig * (q.dv_r * q.dv_r) : 0.0;
    q.dv_phi = va1 - va0;
    q.qp = q.dv_phi < 0.0 ? fac2 * sig * (q.dv_phi * q.dv_phi) : 0.0;
    return q;
}

// energy after the SN dissipation and the temperature floor
// (artificial_viscosity.cpp:187-205, then SetTemperatureFloorCeilValues)
__device__ __forceinline__ double sn_dissipation(const KParams &P, int i, const SnQ &q, double sig,
                                                 double dt, double e)
{
    if (i >= 1 && i < P.nr - 1) {
        const double invdxtheta = ring_tab(P, TAB_INV_DXTHETA, i);  // 1 / (dphi Rmed)
        e = e - dt * q.qr * q.dv_r * cst(P.InvDiffRsup)[i] - dt * q.qp * q.dv_phi * invdxtheta;
    }
    return temperature_floor(P, e, sig);
}

// update_with_artificial_viscosity_SN, artificial_viscosity.cpp:148-211
template <bool SPLIT>
__device__ __forceinline__ void b_sn_q(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(S, FCPT_SIGMA);
    FLD(E, FCPT_ENERGY);
    const bool diss = P.adiabatic && P.artvisc_dissipation;
    const double dt = P.s->dt;
    const int jn = j == ns - 1 ? 0 : j + 1;
    const double vr0 = vr(i, j), vr1 = vr(i + 1, j), va0 = vaz(i, j), va1 = vaz(i, jn);
    const double sig = S(i, j);
    const double e_in = diss ? E(i, j) : 0.0;
    const SnQ q = sn_q_cell(P, vr0, vr1, va0, va1, sig);
    fld<SPLIT>(P, FCPT_Q_R).set(i, j, q.qr);
    fld<SPLIT>(P, FCPT_Q_PHI).set(i, j, q.qp);
    if (diss) E.set(i, j, sn_dissipation(P, i, q, sig, dt, e_in));
}
FCPT_PAIR0(sn_q, 256)

// the SN velocity updates of one cell (artificial_viscosity.cpp:214-248)
template <bool SPLIT = false>
__device__ __forceinline__ double sn_vr_new(const KParams &P, int i, double dt, double v, double s0,
                                            double sm, double qr0, double qrm)
{
    return v - quot<SPLIT>(dt * 2.0, s0 + sm) * (qr0 - qrm) * cst(P.InvDiffRmed)[i];
}

template <bool SPLIT = false>
__device__ __forceinline__ double sn_vphi_new(const KParams &P, int i, double dt, double va, double s0,
                                              double sjp, double qp0, double qpj)
{
    const double invdxtheta = ring_tab(P, TAB_INV_DXTHETA, i);  // 1 / (dphi Rmed)
    return va - quot<SPLIT>(dt * 2.0, s0 + sjp) * (qp0 - qpj) * invdxtheta;
}

template <bool SPLIT>
__device__ __forceinline__ void b_sn_vel(const KParams &P, int bx, int by, int with_viscosity)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(S, FCPT_SIGMA);
    FLD(Qr, FCPT_Q_R);
    FLD(Qp, FCPT_Q_PHI);
    FLD(vaz, FCPT_VAZI);
    FLD(vr, FCPT_VRAD);
    const double dt = P.s->dt;
    const bool inner = i >= 1 && i < P.nr - 1;
    const bool radial = i >= 2 && inner;
    const int jp = j == 0 ? ns - 1 : j - 1;
    const int im = i >= 1 ? i - 1 : 0;
    // in the fused step, recalculate_viscosity (SourceEuler.cpp:205-223), which is
    // pointwise in Sigma and E and follows this kernel, is done here for every ring
    DerivedIn din{0.0, 0.0, 0.0, 0.0};
    const bool visc = with_viscosity && P.adiabatic;  // constant ring profile when isothermal
    if (visc) din = derived_load<SPLIT>(P, i, j, false);
    double s0 = 0.0, sjp = 0.0, sm = 0.0, qp0 = 0.0, qpj = 0.0, qr0 = 0.0, qrm = 0.0;
    double va = 0.0, v = 0.0;
    if (inner) {
        s0 = S(i, j);
        sjp = S(i, jp);
        qp0 = Qp(i, j);
        qpj = Qp(i, jp);
        va = vaz(i, j);
    }
    if (radial) {
        sm = S(im, j);
        qr0 = Qr(i, j);
        qrm = Qr(im, j);
        v = vr(i, j);
    }
    if (visc) derived_store<SPLIT>(P, i, j, din, false, false);
    if (!inner) return;
    if (radial) vr.set(i, j, sn_vr_new<SPLIT>(P, i, dt, v, s0, sm, qr0, qrm));
    vaz.set(i, j, sn_vphi_new<SPLIT>(P, i, dt, va, s0, sjp, qp0, qpj));
}
FCPT_PAIR1(sn_vel, 256, int, with_viscosity)

// recalculate_viscosity, SourceEuler.cpp:205-223
template <bool SPLIT>
__device__ __forceinline__ void b_recalc_viscosity(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr)
    derived_store<SPLIT>(P, i, j, derived_load<SPLIT>(P, i, j, false), false, false, true);
}
FCPT_PAIR0(recalc_viscosity, 256)

// recalculate_derived_disk_quantities, SourceEuler.cpp:225-249
template <bool SPLIT>
__device__ __forceinline__ void b_derived(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr)
    derived_store<SPLIT>(P, i, j, derived_load<SPLIT>(P, i, j, true), true, true);
}
FCPT_PAIR0(derived, 256)


// compute_pressure (SourceEuler.cpp): (gamma - 1) E, or Sigma c_s^2 with the
// isothermal ring profile of c_s
template <bool SPLIT>
__device__ __forceinline__ void b_pressure(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr)
    const double v = P.adiabatic ? (P.gamma - 1.0) * fld<SPLIT>(P, FCPT_ENERGY)(i, j)
                                 : iso_pressure(P, i, fld<SPLIT>(P, FCPT_SIGMA)(i, j));
    fld<SPLIT>(P, FCPT_PRESSURE).set(i, j, v);
}
FCPT_PAIR0(pressure, 256)

// locally isothermal sound speed and scale height from the initial setup
// (compute_sound_speed_normal / compute_scale_height_old, isothermal branches)
template <bool SPLIT>
__device__ __forceinline__ void b_iso_profile(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr)
    const double r = cst(P.Rmed)[i];
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
FCPT_PAIR0(iso_profile, 256)

// the ring-uniform quotients of KParams::ring_tab, each with the expression
// its users had inline (the same bits): 2 / (dphi (Rsup + Rinf)) of the
// azimuthal source term, 1 / (dphi Rmed) of the SN viscosity, 1 / Omega_K
// and Omega_K at Rmed, and 2 / (Rinf(i+1)^2 - Rinf(i)^2) of the viscous v_phi
// update
__global__ void __launch_bounds__(256) k_ring_tables(KParams P, double *tab)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= P.nr) return;
    const size_t w = (size_t)P.nr + 2;
    tab[TAB_SRC_INVDXTHETA * w + i] = 2.0 / (P.dphi * (P.Rsup[i] + P.Rinf[i]));
    const double dxtheta = P.dphi * P.Rmed[i];
    tab[TAB_INV_DXTHETA * w + i] = 1.0 / dxtheta;
    tab[TAB_INV_OMEGA_K * w + i] = 1.0 / omega_kepler(P, P.Rmed[i]);
    tab[TAB_OMEGA_K * w + i] = omega_kepler(P, P.Rmed[i]);
    const double a1 = P.Rinf[i + 1] * P.Rinf[i + 1], a0 = P.Rinf[i] * P.Rinf[i];
    tab[TAB_VISC_AREA * w + i] = 2.0 / (a1 - a0);
}

void launch_ring_tables(const KParams &P, hipStream_t st)
{
    hipLaunchKernelGGL(k_ring_tables, dim3((unsigned)((P.nr + 255) / 256)), dim3(256), 0, st, P, P.ring_tab);
}

// the locally isothermal ring profiles c_s, H and nu = alpha H c_s, with the
// expressions of the 2-D kernels above (k_iso_profile, derived_store)
__global__ void __launch_bounds__(256) k_iso_rings(KParams P)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= P.nr) return;
    const double r = cst(P.Rmed)[i];
    const double vK = sqrt(P.G * P.mstar / r);
    const double h = P.aspect_ratio * pow(r, P.flaring);
    const double c_s = h * vK;
    const double inv_omega_kepler = 1.0 / omega_kepler(P, r);
    const double H = c_s * inv_omega_kepler;
    const double alpha = P.alpha;
    P.iso_cs[i] = c_s;
    P.iso_h[i] = H;
    P.iso_nu[i] = alpha * H * c_s;
}

// ---------------------------------------------------------------------------
// compute_viscous_stress_tensor, viscosity.cpp:139-254: the cell-centred
// divergence and diagonal components at (i, j) from v_r(i), v_r(i + 1),
// v_phi(j), v_phi(j + 1), nu and Sigma of the cell
struct StressDiag {
    double divv, trr, tpp;
};

__device__ __forceinline__ StressDiag stress_diag(const KParams &P, int i, double vr0, double vr1,
                                                  double va0, double va1, double n, double s)
{
    StressDiag d;
    d.divv = (vr1 * cst(P.Rinf)[i + 1] - vr0 * cst(P.Rinf)[i]) * cst(P.InvDiffRsupRb)[i] +
             (va1 - va0) * P.invdphi * cst(P.InvRmed)[i];
    const double drr = (vr1 - vr0) * cst(P.InvDiffRsup)[i];
    d.trr = 2.0 * n * s * (drr - 1.0 / 3.0 * d.divv);
    const double dpp = (va1 - va0) * P.invdphi * cst(P.InvRmed)[i] + 0.5 * (vr1 + vr0) * cst(P.InvRmed)[i];
    d.tpp = 2.0 * n * s * (dpp - 1.0 / 3.0 * d.divv);
    return d;
}

// the corner component tau_r_phi at (i >= 1, j) from v_phi(i), v_phi(i - 1),
// v_r(j), v_r(j - 1) and nu, Sigma of the four cells around the corner
__device__ __forceinline__ double stress_rp(const KParams &P, int i, double va0, double vam, double vr0,
                                            double vrp, double n, double nm, double njp, double nmp,
                                            double s, double sm, double sjp, double smp)
{
    const double dvazirdr = (va0 * cst(P.InvRmed)[i] - vam * cst(P.InvRmed)[i - 1]) * cst(P.InvDiffRmed)[i];
    const double dvrdphi = (vr0 - vrp) * P.invdphi;
    const double drp = cst(P.Rinf)[i] * dvazirdr + dvrdphi * cst(P.InvRinf)[i];
    const double n4 = 0.25 * (n + nm + njp + nmp);
    const double s4 = 0.25 * (s + sm + sjp + smp);
    return n4 * s4 * drp;
}

template <bool SPLIT>
__device__ __forceinline__ void b_stress(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr)
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(nu, FCPT_VISCOSITY);
    FLD(S, FCPT_SIGMA);
    const int jn = j == ns - 1 ? 0 : j + 1;
    const int jp = j == 0 ? ns - 1 : j - 1;
    const int im = i >= 1 ? i - 1 : 0;
    const double vr0 = vr(i, j), vr1 = vr(i + 1, j), va0 = vaz(i, j), va1 = vaz(i, jn);
    const bool iso = !P.adiabatic;  // nu is a ring profile
    const double n = iso ? cst(P.iso_nu)[i] : nu(i, j), s = S(i, j);
    const double vam = vaz(im, j), vrp = vr(i, jp);
    const double nm = iso ? cst(P.iso_nu)[im] : nu(im, j);
    const double njp = iso ? n : nu(i, jp), nmp = iso ? nm : nu(im, jp);
    const double sm = S(im, j), sjp = S(i, jp), smp = S(im, jp);
    const StressDiag d = stress_diag(P, i, vr0, vr1, va0, va1, n, s);
    fld<SPLIT>(P, FCPT_DIV_V).set(i, j, d.divv);
    fld<SPLIT>(P, FCPT_TAU_R_R).set(i, j, d.trr);
    fld<SPLIT>(P, FCPT_TAU_PHI_PHI).set(i, j, d.tpp);
    if (i >= 1)
        fld<SPLIT>(P, FCPT_TAU_R_PHI).set(i, j, stress_rp(P, i, va0, vam, vr0, vrp, n, nm, njp, nmp, s, sm, sjp, smp));
}
FCPT_PAIR0(stress, 256)

// update_velocities_with_viscosity, viscosity.cpp:355-426, for one cell
template <bool SPLIT = false>
__device__ __forceinline__ double visc_vphi_new(const KParams &P, int i, double dt, double va, double s0,
                                                double sjp, double trp0, double trp1, double tpp0,
                                                double tppj)
{
    const double sigma_avg = 0.5 * (s0 + sjp);
    const double a1 = cst(P.Rinf)[i + 1] * cst(P.Rinf)[i + 1], a0 = cst(P.Rinf)[i] * cst(P.Rinf)[i];
    const double dVp = quot<SPLIT>(dt * cst(P.InvRmed)[i], sigma_avg) *
                       (ring_tab(P, TAB_VISC_AREA, i) * (a1 * trp1 - a0 * trp0) + (tpp0 - tppj) * P.invdphi);
    return va + dVp;
}

template <bool SPLIT = false>
__device__ __forceinline__ double visc_vr_new(const KParams &P, int i, double dt, double v, double s0,
                                              double sm, double trr0, double trrm, double trp0,
                                              double trpn, double tpp0, double tppm)
{
    const double sigma_avg = 0.5 * (s0 + sm);
    const double dVr = quot<SPLIT>(quot<SPLIT>(dt, sigma_avg) * P.radial_viscosity_factor * 2.0,
                                   cst(P.Rmed)[i] + cst(P.Rmed)[i - 1]) *
                       ((cst(P.Rmed)[i] * trr0 - cst(P.Rmed)[i - 1] * trrm) * cst(P.InvDiffRmed)[i] +
                        (trpn - trp0) * P.invdphi - 0.5 * (tpp0 + tppm));
    return v + dVr;
}

template <bool SPLIT>
__device__ __forceinline__ void b_visc_update(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(1, P.nr - 1)
    FLD(S, FCPT_SIGMA);
    FLD(trr, FCPT_TAU_R_R);
    FLD(tpp, FCPT_TAU_PHI_PHI);
    FLD(trp, FCPT_TAU_R_PHI);
    FLD(vaz, FCPT_VAZI);
    FLD(vr, FCPT_VRAD);
    const double dt = P.s->dt;
    const int jp = j == 0 ? ns - 1 : j - 1;
    const int jn = j == ns - 1 ? 0 : j + 1;
    const double s0 = S(i, j), sjp = S(i, jp), sm = S(i - 1, j);
    const double trp0 = trp(i, j), trp1 = trp(i + 1, j), trpn = trp(i, jn);
    const double tpp0 = tpp(i, j), tppj = tpp(i, jp), tppm = tpp(i - 1, j);
    const double trr0 = trr(i, j), trrm = trr(i - 1, j);
    const double va = vaz(i, j), v = vr(i, j);
    vaz.set(i, j, visc_vphi_new<SPLIT>(P, i, dt, va, s0, sjp, trp0, trp1, tpp0, tppj));
    if (i >= 2) vr.set(i, j, visc_vr_new<SPLIT>(P, i, dt, v, s0, sm, trr0, trrm, trp0, trpn, tpp0, tppm));
}
FCPT_PAIR0(visc_update, 256)

// ---------------------------------------------------------------------------
// Fused kernels of the step.  Each reads the velocities of its view (P.f[VRAD],
// P.f[VAZI]) and writes the velocities of every row, updated or passed
// through, to the view's partner buffers P.valt: a cell's neighbours then
// always see the old values, exactly as between the separate kernels above,
// whose per-cell expressions they call, and whose intermediate tensors they
// recompute where a neighbour needs them instead of storing and reloading
// them.  The step alternates two views with swapped velocity pointers; the
// transport's velocity kernel writes the final velocities to the fields.
template <bool SPLIT>
__device__ __forceinline__ Fld<SPLIT> alt_vr(const KParams &P)
{
    return buf<SPLIT>(P, P.valt[0], P.bg[FCPT_VRAD], FCPT_VRAD);
}
template <bool SPLIT>
__device__ __forceinline__ Fld<SPLIT> alt_vaz(const KParams &P)
{
    return buf<SPLIT>(P, P.valt[1], P.bg[FCPT_VAZI], FCPT_VAZI);
}

// update_with_sourceterms' two momentum updates (SourceEuler.cpp:325-428) in
// one pass over rows 0 .. nr; v_r is updated on rings 2 .. nr-2, v_phi on
// 1 .. nr-2.  A locally isothermal disk's pressure is formed from Sigma.
template <bool SPLIT>
__device__ __forceinline__ void b_src_fused(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr + 1)
    const int nr = P.nr;
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    const Fld<SPLIT> vr_o = alt_vr<SPLIT>(P);
    const Fld<SPLIT> vaz_o = alt_vaz<SPLIT>(P);
    const bool upd_p = i >= 1 && i < nr - 1;
    const bool upd_r = i >= 2 && upd_p;
    if (!upd_p) {
        vr_o.set_raw(i, j, vr.raw(i, j));
        if (i < nr) vaz_o.set_raw(i, j, vaz.raw(i, j));
        return;
    }
    FLD(S, FCPT_SIGMA);
    FLD(Pr, FCPT_PRESSURE);
    const Fld<false> Phi = phi_fld(P);
    const bool iso = !P.adiabatic;
    const double dt = P.s->dt;
    const int jn = j == ns - 1 ? 0 : j + 1;
    const int jp = j == 0 ? ns - 1 : j - 1;
    const double s0 = S(i, j), sjp = S(i, jp);
    const double p0 = iso ? iso_pressure(P, i, s0) : Pr(i, j);
    const double pjp = iso ? iso_pressure(P, i, sjp) : Pr(i, jp);
    const double phi0 = Phi(i, j), phijp = Phi(i, jp);
    const double va0 = vaz(i, j);
    double v = 0.0, sm = 0.0, pm = 0.0, phim = 0.0, va1 = 0.0, vam0 = 0.0, vam1 = 0.0;
    if (upd_r) {
        v = vr(i, j);
        sm = S(i - 1, j);
        pm = iso ? iso_pressure(P, i - 1, sm) : Pr(i - 1, j);
        phim = Phi(i - 1, j);
        va1 = vaz(i, jn);
        vam0 = vaz(i - 1, j);
        vam1 = vaz(i - 1, jn);
    }
    const typename Fld<SPLIT>::T vr_raw = vr.raw(i, j);
    vaz_o.set(i, j, src_vphi_new<SPLIT>(P, i, dt, va0, s0, sjp, p0, pjp, phi0, phijp));
    if (upd_r) vr_o.set(i, j, src_vr_new<SPLIT>(P, i, dt, v, s0, sm, p0, pm, phi0, phim, va0, va1, vam0, vam1));
    else vr_o.set_raw(i, j, vr_raw);
}
FCPT_PAIR0(src_fused, 256)

// the TW or SN artificial viscosity (artificial_viscosity.cpp:35-248) in one
// pass: the tensor of the cell, of its azimuthal neighbour j - 1 and of its
// radial neighbour i - 1, the dissipation and temperature floor of E, the
// viscosity recalculation when asked, and the velocity updates
template <bool SPLIT>
__device__ __forceinline__ void b_av_fused(const KParams &P, int bx, int by, int with_viscosity)
{
    CELL_LOOP_HEAD(0, P.nr + 1)
    const int nr = P.nr;
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    const Fld<SPLIT> vr_o = alt_vr<SPLIT>(P);
    const Fld<SPLIT> vaz_o = alt_vaz<SPLIT>(P);
    if (i == nr) {
        vr_o.set_raw(i, j, vr.raw(i, j));
        return;
    }
    FLD(S, FCPT_SIGMA);
    FLD(E, FCPT_ENERGY);
    const bool sn = P.artvisc == FCPT_ARTVISC_SN;
    const bool diss = P.adiabatic && P.artvisc_dissipation;
    const bool visc = with_viscosity && P.adiabatic;  // constant ring profile when isothermal
    const double dt = P.s->dt;
    const int jn = j == ns - 1 ? 0 : j + 1;
    const int jp = j == 0 ? ns - 1 : j - 1;
    const bool inner = i >= 1 && i < nr - 1;
    const bool radial = i >= 2 && inner;
    const int im = i >= 1 ? i - 1 : 0;
    // rows that are not updated are copied as stored (in split storage a
    // decode / encode round trip would round the deviation)
    const typename Fld<SPLIT>::T vr_raw = vr.raw(i, j), va_raw = vaz.raw(i, j);
    const double vr0 = vr.dec(i, vr_raw), vr1 = vr(i + 1, j), va0 = vaz.dec(i, va_raw), va1 = vaz(i, jn);
    const double sig = S(i, j);
    const double e_in = (diss || visc) ? E(i, j) : 0.0;
    double vrj0 = 0.0, vrj1 = 0.0, vaj0 = 0.0, sjp = 0.0;  // cell (i, j - 1)
    if (inner) {
        vrj0 = vr(i, jp);
        vrj1 = vr(i + 1, jp);
        vaj0 = vaz(i, jp);
        sjp = S(i, jp);
    }
    double vrm0 = 0.0, vam0 = 0.0, vam1 = 0.0, sm = 0.0;  // cell (i - 1, j)
    if (radial) {
        vrm0 = vr(im, j);
        vam0 = vaz(im, j);
        vam1 = vaz(im, jn);
        sm = S(im, j);
    }
    double qr0, qp0, qpj = 0.0, qrm = 0.0, qpm = 0.0, e = e_in;
    if (sn) {
        const SnQ q = sn_q_cell(P, vr0, vr1, va0, va1, sig);
        qr0 = q.qr;
        qp0 = q.qp;
        if (diss) e = sn_dissipation(P, i, q, sig, dt, e_in);
        if (inner) qpj = sn_q_cell(P, vrj0, vrj1, vaj0, va0, sjp).qp;
        if (radial) qrm = sn_q_cell(P, vrm0, vr0, vam0, vam1, sm).qr;
    } else {
        const TwQ q = tw_q_cell(P, i, vr0, vr1, va0, va1, sig);
        qr0 = q.qr;
        qp0 = q.qp;
        if (diss) e = tw_dissipation(P, i, q, sig, dt, e_in);
        if (inner) qpj = tw_q_cell(P, i, vrj0, vrj1, vaj0, va0, sjp).qp;
        if (radial) {
            const TwQ m = tw_q_cell(P, im, vrm0, vr0, vam0, vam1, sm);
            qrm = m.qr;
            qpm = m.qp;
        }
    }
    if (diss) e = E.set_get(i, j, e);
    if (visc) derived_store<SPLIT>(P, i, j, DerivedIn{e, sig, 0.0, 0.0}, false, false);
    if (inner)
        vaz_o.set(i, j, sn ? sn_vphi_new<SPLIT>(P, i, dt, va0, sig, sjp, qp0, qpj)
                           : tw_vphi_new<SPLIT>(P, i, dt, va0, sig, sjp, qp0, qpj));
    else
        vaz_o.set_raw(i, j, va_raw);
    if (radial)
        vr_o.set(i, j, sn ? sn_vr_new<SPLIT>(P, i, dt, vr0, sig, sm, qr0, qrm)
                          : tw_vr_new<SPLIT>(P, i, dt, vr0, sig, sm, qr0, qrm, qp0, qpm));
    else
        vr_o.set_raw(i, j, vr_raw);
}
FCPT_PAIR1(av_fused, 256, int, with_viscosity)

// The kick's source terms and artificial viscosity as one radial walk:
// update_with_sourceterms, compression_heating (adiabatic) and the TW or SN
// artificial viscosity with its dissipation and recalculate_viscosity, the
// per-cell steps of b_src_fused, b_src_energy and b_av_fused in that order,
// after the potential pass.  A wavefront owns kKickSegW consecutive sectors
// and `tile` consecutive rings: lane k holds sector j0 - 2 + k (lanes 0, 1,
// 62 and 63 are the halo) and walks the rings outwards.  Each ring is loaded
// once (v_r, v_phi, Sigma, the potential, and the pressure of an adiabatic
// disk); the source-term velocities of a ring are formed one ring ahead of the
// artificial viscosity that reads them, the neighbouring sectors' values come
// from the adjacent lanes by DPP moves (each stage spoils one more lane at
// either edge), and the tensor of ring i - 1 carries over.  E is read and
// written only at the lane's own cells, so no wavefront reads what another
// writes, and the derived fields it stores (c_s, H, nu) are not read here.
// Every value goes through the helpers
// of the separate kernels with the same arguments, so FP64 storage gives the
// same fields bit for bit; in split storage the source-term velocities stay
// in FP64 between the stages.
constexpr int kKickSegW = 60;

struct KickRing {
    double vr, va, s, p, phi, e;  // the lane's sector (p: pressure, phi: potential, e: E)
    double sL, vaR;               // Sigma of sector j - 1, v_phi of sector j + 1
};

template <bool SPLIT, bool ADI, bool SN>
__device__ __forceinline__ void b_kick_walk(const KParams &P, int bx, int by, int tile)
{
    const int ns = P.ns, nr = P.nr;
    const int k = (int)threadIdx.x;  // lane
    const int i0 = __builtin_amdgcn_readfirstlane((by * (int)blockDim.y + (int)threadIdx.y) * tile);
    if (i0 >= nr) return;
    const int i1 = i0 + tile < nr ? i0 + tile : nr;  // rings [i0, i1)
    const int j0 = bx * kKickSegW;
    const int Ls = ns - j0 < kKickSegW ? ns - j0 : kKickSegW;
    const int j = (j0 - 2 + k + 2 * ns) % ns;
    const bool own = k >= 2 && k < 2 + Ls;  // lanes that write
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(S, FCPT_SIGMA);
    FLD(E, FCPT_ENERGY);
    const Fld<SPLIT> vr_o = alt_vr<SPLIT>(P);
    const Fld<SPLIT> vaz_o = alt_vaz<SPLIT>(P);
    const double dt = P.s->dt;
    constexpr bool sn = SN;
    const bool diss = ADI && P.artvisc_dissipation;
    auto load = [&](int r) {
        KickRing w;
        w.vr = vr(r, j);
        if (r < nr) {
            w.va = vaz(r, j);
            w.s = S(r, j);
            // P = (gamma - 1) E as the last step stored it (compute_pressure):
            // with the partner buffer no wavefront writes E during the walk
            w.e = ADI ? E(r, j) : 0.0;
            w.p = ADI ? (P.e_partner ? (P.gamma - 1.0) * w.e : fld<SPLIT>(P, FCPT_PRESSURE)(r, j))
                      : iso_pressure(P, r, w.s);
            w.phi = phi_fld(P)(r, j);
        } else {
            w.va = w.s = w.p = w.phi = w.e = 0.0;
        }
        w.sL = from_left(w.s);
        w.vaR = from_right(w.va);
        return w;
    };
    // the source-term velocities of ring r (b) from ring r - 1 (a); rows
    // that are not updated keep their values
    auto src = [&](int r, const KickRing &a, const KickRing &b, double &v1r, double &v1a) {
        const double pjp = from_left(b.p), phijp = from_left(b.phi);
        const bool upd_p = r >= 1 && r < nr - 1;
        const bool upd_r = r >= 2 && upd_p;
        v1a = upd_p ? src_vphi_new<SPLIT>(P, r, dt, b.va, b.s, b.sL, b.p, pjp, b.phi, phijp) : b.va;
        v1r = upd_r ? src_vr_new<SPLIT>(P, r, dt, b.vr, b.s, a.s, b.p, a.p, b.phi, a.phi, b.va, b.vaR, a.va, a.vaR)
                    : b.vr;
    };
    // the tensor of cell (r, j) from the source-term velocities
    struct Q {
        TwQ tw;
        SnQ sn;
        double qr, qp;
    };
    auto qcell = [&](int r, double vr0, double vr1, double va0, double va1, double sig) {
        Q q;
        if constexpr (SN) {
            q.sn = sn_q_cell(P, vr0, vr1, va0, va1, sig);
            q.qr = q.sn.qr;
            q.qp = q.sn.qp;
        } else {
            q.tw = tw_q_cell(P, r, vr0, vr1, va0, va1, sig);
            q.qr = q.tw.qr;
            q.qp = q.tw.qp;
        }
        return q;
    };
    KickRing R0, R1;  // rings i, i + 1
    double v1r0 = 0.0, v1a0 = 0.0;  // source-term velocities of ring i
    double qr_m = 0.0, qp_m = 0.0, s_m = 0.0;  // tensor and Sigma of cell (i - 1, j)
    if (i0 >= 1) {
        const KickRing Rm2 = load(i0 >= 2 ? i0 - 2 : 0);  // unused by ring 0
        const KickRing Rm = load(i0 - 1);
        R0 = load(i0);
        double v1rm, v1am;
        src(i0 - 1, Rm2, Rm, v1rm, v1am);
        src(i0, Rm, R0, v1r0, v1a0);
        const Q qm = qcell(i0 - 1, v1rm, v1r0, v1am, from_right(v1am), Rm.s);
        qr_m = qm.qr;
        qp_m = qm.qp;
        s_m = Rm.s;
    } else {
        R0 = load(i0);
        src(i0, R0, R0, v1r0, v1a0);  // ring 0 is passed through
    }
    for (int i = i0; i < i1; ++i) {
        R1 = load(i + 1);  // ring nr holds only v_r
        double v1r1, v1a1;
        src(i + 1, R0, R1, v1r1, v1a1);
        const double v1aR = from_right(v1a0);  // v_phi of sector j + 1
        const Q q = qcell(i, v1r0, v1r1, v1a0, v1aR, R0.s);
        const double qpj = from_left(q.qp);  // tensor of cell (i, j - 1)
        const bool inner = i >= 1 && i < nr - 1;
        const bool radial = i >= 2 && inner;
        if (own) {
            if constexpr (ADI) {
                // compression_heating (rings 0 .. nr-2), then the dissipation;
                // the result to E or (e_partner) to its partner buffer, every cell
                const Fld<SPLIT> Eo = P.e_partner ? buf<SPLIT>(P, P.ealt, P.bg[FCPT_ENERGY], FCPT_ENERGY) : E;
                double e = R0.e;
                if (i < nr - 1) e = E.stored(i, j, src_energy_new(P, i, dt, e, v1r0, v1r1, v1a0, v1aR));
                if (diss) e = sn ? sn_dissipation(P, i, q.sn, R0.s, dt, e) : tw_dissipation(P, i, q.tw, R0.s, dt, e);
                if (i < nr - 1 || diss) {
                    e = Eo.set_get(i, j, e);  // the partner has E's background and key
                } else if (P.e_partner) {
                    Eo.set_raw(i, j, E.raw(i, j));
                }
                // recalculate_viscosity
                const AdiDerived d = adi_derived<SPLIT>(P, i, e, R0.s);
                if (!lazy_derived<SPLIT>(P)) {
                    fld<SPLIT>(P, FCPT_SOUNDSPEED).set(i, j, d.cs);
                    fld<SPLIT>(P, FCPT_SCALE_HEIGHT).set(i, j, d.h);
                }
                fld<SPLIT>(P, FCPT_VISCOSITY).set(i, j, d.nu);
            }
            if (inner)
                vaz_o.set(i, j, sn ? sn_vphi_new<SPLIT>(P, i, dt, v1a0, R0.s, R0.sL, q.qp, qpj)
                                   : tw_vphi_new<SPLIT>(P, i, dt, v1a0, R0.s, R0.sL, q.qp, qpj));
            else
                vaz_o.set_raw(i, j, vaz.raw(i, j));
            if (radial)
                vr_o.set(i, j, sn ? sn_vr_new<SPLIT>(P, i, dt, v1r0, R0.s, s_m, q.qr, qr_m)
                                  : tw_vr_new<SPLIT>(P, i, dt, v1r0, R0.s, s_m, q.qr, qr_m, q.qp, qp_m));
            else
                vr_o.set_raw(i, j, vr.raw(i, j));
        }
        qr_m = q.qr;
        qp_m = q.qp;
        s_m = R0.s;
        R0 = R1;
        v1r0 = v1r1;
        v1a0 = v1a1;
    }
    if (i1 == nr && own) vr_o.set_raw(nr, j, vr.raw(nr, j));  // the outer v_r row, untouched
}
template <bool SPLIT, bool ADI, bool SN>
__global__ void __launch_bounds__(256) k_kick_walk(KParams P, int tile)
{
    const BlockXY bb = xcd_block();
    b_kick_walk<SPLIT, ADI, SN>(P, bb.x, bb.y, tile);
}
// batched form: each disk's tile is its resolved kick_tile
template <bool SPLIT, bool ADI, bool SN>
__global__ void __launch_bounds__(256) kb_kick_walk(const KParams *Ps, const BatchEntry *map)
{
    const BatchEntry e = batch_entry(map);
    const KParams &P = batch_disk(Ps, e.disk);
    b_kick_walk<SPLIT, ADI, SN>(P, (int)(e.yx & 0xffffu), (int)(e.yx >> 16), P.kick_tile);
}

// compute_viscous_stress_tensor and update_velocities_with_viscosity
// (viscosity.cpp:139-426) in one pass: the stress components a cell's update
// needs -- tau_rr at (i, j), (i - 1, j); tau_pp at (i, j), (i, j - 1),
// (i - 1, j); tau_rp at (i, j), (i + 1, j), (i, j + 1) -- from the velocities,
// nu and Sigma around it.  The tensor and div v of the cell are stored when a
// later kernel of the step reads them (the viscous heating of SubStep3 in an
// adiabatic disk).
template <bool SPLIT>
__device__ __forceinline__ void b_visc_fused(const KParams &P, int bx, int by)
{
    const bool store_tensor = P.adiabatic && P.heating_viscous;
    CELL_LOOP_HEAD(0, P.nr + 1)
    const int nr = P.nr;
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    const Fld<SPLIT> vr_o = alt_vr<SPLIT>(P);
    const Fld<SPLIT> vaz_o = alt_vaz<SPLIT>(P);
    if (i == nr) {
        vr_o.set_raw(i, j, vr.raw(i, j));
        return;
    }
    FLD(nu, FCPT_VISCOSITY);
    FLD(S, FCPT_SIGMA);
    const bool iso = !P.adiabatic;  // nu is a ring profile
    const double dt = P.s->dt;
    const int jn = j == ns - 1 ? 0 : j + 1;
    const int jp = j == 0 ? ns - 1 : j - 1;
    const int im = i >= 1 ? i - 1 : 0;
    const bool upd_p = i >= 1 && i < nr - 1;
    const bool upd_r = i >= 2 && upd_p;
    const int ip = upd_p ? i + 1 : i;
    // the cell and the corner (i, j); rows that are not updated are copied as stored
    const typename Fld<SPLIT>::T vr_raw = vr.raw(i, j), va_raw = vaz.raw(i, j);
    const double vr0 = vr.dec(i, vr_raw), vr1 = vr(i + 1, j), va0 = vaz.dec(i, va_raw), va1 = vaz(i, jn);
    const double n = iso ? cst(P.iso_nu)[i] : nu(i, j), s = S(i, j);
    const double vam = vaz(im, j), vrp = vr(i, jp);
    const double nm = iso ? cst(P.iso_nu)[im] : nu(im, j);
    const double njp = iso ? n : nu(i, jp), nmp = iso ? nm : nu(im, jp);
    const double sm = S(im, j), sjp = S(i, jp), smp = S(im, jp);
    // corner (i + 1, j) and cell (i, j - 1): the v_phi update
    double va_p = 0.0, vr1p = 0.0, np_ = 0.0, npj = 0.0, sp = 0.0, spj = 0.0, vaj = 0.0;
    if (upd_p) {
        va_p = vaz(ip, j);
        vr1p = vr(ip, jp);
        np_ = iso ? cst(P.iso_nu)[ip] : nu(ip, j);
        npj = iso ? np_ : nu(ip, jp);
        sp = S(ip, j);
        spj = S(ip, jp);
        vaj = vaz(i, jp);
    }
    // cell (i - 1, j) and corner (i, j + 1): the v_r update
    double vrm = 0.0, vam1 = 0.0, vrn = 0.0, nn = 0.0, nmn = 0.0, sn = 0.0, smn = 0.0;
    if (upd_r) {
        vrm = vr(im, j);
        vam1 = vaz(im, jn);
        vrn = vr(i, jn);
        nn = iso ? n : nu(i, jn);
        nmn = iso ? nm : nu(im, jn);
        sn = S(i, jn);
        smn = S(im, jn);
    }
    const StressDiag d = stress_diag(P, i, vr0, vr1, va0, va1, n, s);
    const double trp0 = i >= 1 ? stress_rp(P, i, va0, vam, vr0, vrp, n, nm, njp, nmp, s, sm, sjp, smp) : 0.0;
    if (store_tensor) {
        fld<SPLIT>(P, FCPT_DIV_V).set(i, j, d.divv);
        fld<SPLIT>(P, FCPT_TAU_R_R).set(i, j, d.trr);
        fld<SPLIT>(P, FCPT_TAU_PHI_PHI).set(i, j, d.tpp);
        if (i >= 1) fld<SPLIT>(P, FCPT_TAU_R_PHI).set(i, j, trp0);
    }
    if (upd_p) {
        const double trp1 = stress_rp(P, ip, va_p, va0, vr1, vr1p, np_, n, npj, njp, sp, s, spj, sjp);
        const double tppj = stress_diag(P, i, vrp, vr1p, vaj, va0, njp, sjp).tpp;
        vaz_o.set(i, j, visc_vphi_new<SPLIT>(P, i, dt, va0, s, sjp, trp0, trp1, d.tpp, tppj));
    } else {
        vaz_o.set_raw(i, j, va_raw);
    }
    if (upd_r) {
        const StressDiag dm = stress_diag(P, im, vrm, vr0, vam, vam1, nm, sm);
        const double trpn = stress_rp(P, i, va1, vam1, vrn, vr0, nn, nmn, n, nm, sn, smn, s, sm);
        vr_o.set(i, j, visc_vr_new<SPLIT>(P, i, dt, vr0, s, sm, d.trr, dm.trr, trp0, trpn, d.tpp, dm.tpp));
    } else {
        vr_o.set_raw(i, j, vr_raw);
    }
}
FCPT_PAIR0(visc_fused, 256)

// the velocities of view P copied as stored to its partner buffers (a LeapFrog
// step whose second kick ends on the swapped view hands them back this way)
template <bool SPLIT>
__device__ __forceinline__ void b_copy_velocities(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr + 1)
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    alt_vr<SPLIT>(P).set_raw(i, j, vr.raw(i, j));
    if (i < P.nr) alt_vaz<SPLIT>(P).set_raw(i, j, vaz.raw(i, j));
}
FCPT_PAIR0(copy_velocities, 256)

// ---------------------------------------------------------------------------
// Q- of the beta cooling (thermal_relaxation, SourceEuler.cpp:632-690) for
// one active cell
template <bool SPLIT>
__device__ __forceinline__ double cooling_beta_cell(const KParams &P, int i, int j, double time, double e,
                                                    double sig)
{
    if (!P.cooling_beta_enabled) return 0.0;
    const double omega_k = ring_tab(P, TAB_OMEGA_K, i);  // Omega_K(Rmed)
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
        delta_E -= energy_floor<SPLIT>(P, sig);
    return 0.0 + delta_E * omega_k * beta_inv;
}

// Q+ of the viscous heating (viscous_heating, SourceEuler.cpp:496-536) from
// the cell's stress tensor; tau_r_phi is the mean of the cell's four corners
template <bool SPLIT = false>
__device__ __forceinline__ double viscous_heating_q(const KParams &P, double nu, double sig, double trr,
                                                    double tpp, double dv, double tau_r_phi)
{
    double q = quot<SPLIT>(1.0, 2.0 * nu * sig) * (trr * trr + 2 * (tau_r_phi * tau_r_phi) + tpp * tpp);
    q += (2.0 / 9.0) * nu * sig * (dv * dv);
    q *= P.heating_viscous_factor;
    return 0.0 + q;
}

// Q- and Q+ for one active cell, the tensor read from its fields
template <bool SPLIT>
__device__ __forceinline__ void heating_cooling_cell(const KParams &P, int i, int j, double time,
                                                     double e, double sig, double &qplus,
                                                     double &qminus)
{
    const int ns = P.ns;
    qminus = cooling_beta_cell<SPLIT>(P, i, j, time, e, sig);
    qplus = 0.0;
    if (P.heating_viscous) {
        const double nu = fld<SPLIT>(P, FCPT_VISCOSITY)(i, j);
        if (nu != 0.0) {
            FLD(trp, FCPT_TAU_R_PHI);
            const int jn = j == ns - 1 ? 0 : j + 1;
            const double tau_r_phi = 0.25 * (trp(i, j) + trp(i + 1, j) + trp(i, jn) + trp(i + 1, jn));
            qplus = viscous_heating_q<SPLIT>(P, nu, sig, fld<SPLIT>(P, FCPT_TAU_R_R)(i, j),
                                      fld<SPLIT>(P, FCPT_TAU_PHI_PHI)(i, j), fld<SPLIT>(P, FCPT_DIV_V)(i, j),
                                      tau_r_phi);
        }
    }
}


// the radiative correction of SubStep3 (SourceEuler.cpp:920-924; its
// std::pow(., 4) and std::pow(., 3) by compensated powering)
template <bool SPLIT = false>
__device__ __forceinline__ double rad_alpha(const KParams &P, double h, double sig, double e)
{
    const double inv_pow4 = pow4_comp(quot<SPLIT>(P.mu * (P.gamma - 1.0), P.Rgas * sig));
    return 1.0 + 2.0 * h * 4.0 * P.sigma_sb / P.c_light * inv_pow4 * pow3_comp(e);
}

// SubStep3 (SourceEuler.cpp:859-954) at one cell from its Q+ and Q- before
// the radiative correction; writes T, Q+, Q- and E
template <bool SPLIT>
__device__ __forceinline__ void substep3_cell(const KParams &P, int i, int j, double e, double sig, double h,
                                              double qplus, double qminus)
{
    const double dt = P.s->dt;
    // compute_temperature (SourceEuler.cpp:1378-1408) before the update
    const double c_v_inv = P.mu / P.Rgas * (P.gamma - 1.0);
    const double temperature = c_v_inv * e / sig;
    if (i >= 1 && i < P.nr - 1) {
        const double alpha = rad_alpha<SPLIT>(P, h, sig, e);
        if constexpr (SPLIT) {
            const double ra = recip<true>(alpha);
            qplus *= ra;
            qminus *= ra;
        } else {  // the same bits through one reciprocal
            const SharedDen A = shared_den(alpha);
            qplus = shared_div(qplus, A);
            qminus = shared_div(qminus, A);
        }
        double energy_new = e + dt * (qplus - qminus);
        const double SigmaFloor = 10.0 * P.sigma0 * P.sigma_floor;
        if ((sig < SigmaFloor)) {
            // TAU_EFF is never computed without surface cooling: it holds 0
            const double tau_eff = 0.0;
            const double e4 = qplus * tau_eff / (2.0 * P.sigma_sb);
            const double constant = (P.Rgas / P.mu * sig / (P.gamma - 1.0));
            energy_new = sqrt(sqrt(e4)) * constant;  // pow(e4, 1 / 4): 0 here, as tau_eff is 0
            qminus = qplus;
        }
        e = energy_new;
    } else {
        qplus = qminus = 0.0;
    }
    if (!lazy_derived<SPLIT>(P)) fld<SPLIT>(P, FCPT_TEMPERATURE).set(i, j, temperature);
    fld<SPLIT>(P, FCPT_QPLUS).set(i, j, qplus);
    fld<SPLIT>(P, FCPT_QMINUS).set(i, j, qminus);
    fld<SPLIT>(P, FCPT_ENERGY).set(i, j, temperature_floor<SPLIT>(P, e, sig));
}

template <bool SPLIT>
__device__ __forceinline__ void b_substep3(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr)
    const double sig = fld<SPLIT>(P, FCPT_SIGMA)(i, j);
    const double e = fld<SPLIT>(P, FCPT_ENERGY)(i, j);
    const double h = fld<SPLIT>(P, FCPT_SCALE_HEIGHT)(i, j);
    double qplus = 0.0, qminus = 0.0;
    if (i >= 1 && i < P.nr - 1) heating_cooling_cell<SPLIT>(P, i, j, P.s->time, e, sig, qplus, qminus);
    substep3_cell<SPLIT>(P, i, j, e, sig, h, qplus, qminus);
}
FCPT_PAIR0(substep3, 256)

// compute_heating_cooling_for_CFL, SourceEuler.cpp:1410-1450 (after update_viscosity
// and the stress tensor)
template <bool SPLIT>
__device__ __forceinline__ void b_heatcool_cfl(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr)
    double qplus = 0.0, qminus = 0.0;
    if (i >= 1 && i < P.nr - 1) {
        const double e = fld<SPLIT>(P, FCPT_ENERGY)(i, j);
        const double sig = fld<SPLIT>(P, FCPT_SIGMA)(i, j);
        heating_cooling_cell<SPLIT>(P, i, j, P.s->time, e, sig, qplus, qminus);
        const double alpha = rad_alpha<SPLIT>(P, fld<SPLIT>(P, FCPT_SCALE_HEIGHT)(i, j), sig, e);
        qplus /= alpha;
        qminus /= alpha;
    }
    fld<SPLIT>(P, FCPT_QPLUS).set(i, j, qplus);
    fld<SPLIT>(P, FCPT_QMINUS).set(i, j, qminus);
}
FCPT_PAIR0(heatcool_cfl, 256)

// The fused viscous pass as a radial walk.  A wavefront owns kViscSegW
// consecutive sectors and `tile` consecutive rings: lane k holds sector
// j0 - 1 + k (lanes 0 and 63 are the halo) and walks the rings outwards.
// Each ring is loaded once (v_r, v_phi, Sigma, nu of the lane's own sector;
// the ring after next is prefetched before the current ring's stores), the
// neighbouring sectors' values come from the adjacent lanes by DPP moves,
// and every stress component is formed once per cell: the diagonal of ring
// i, whose tau_phi_phi at sector j - 1 is the left lane's, and tau_rp of
// corner (i + 1, j), whose value at sector j + 1 is the right lane's; the
// diagonal of ring i - 1 and the corners of ring i carry over from the
// previous step.  Every value goes through the helpers b_visc_fused calls,
// with the same arguments, so the two are bit-identical.  With ENERGY the walk
// also runs SubStep3 on each cell it updates, its viscous heating formed from
// the tensor in registers (the corner mean takes tau_rp of corner (i + 1,
// j + 1) from the right lane), exactly as b_substep3 forms it from the stored
// tensor, which the walk then does not store.
constexpr int kViscSegW = 62;  // sectors written per wavefront

struct ViscRing {
    double vr, va, s, n;  // the lane's sector
    double vrL, sL, nL;   // sector j - 1
    double vaR;           // v_phi of sector j + 1
};

template <bool SPLIT, bool ENERGY>
__device__ __forceinline__ void b_visc_walk(const KParams &P, int bx, int by, int tile)
{
    const int ns = P.ns, nr = P.nr;
    const int k = (int)threadIdx.x;  // lane
    const int i0 = __builtin_amdgcn_readfirstlane((by * (int)blockDim.y + (int)threadIdx.y) * tile);
    if (i0 >= nr) return;
    const int i1 = i0 + tile < nr ? i0 + tile : nr;  // rings [i0, i1)
    const int j0 = bx * kViscSegW;
    const int Ls = ns - j0 < kViscSegW ? ns - j0 : kViscSegW;
    const int j = (j0 - 1 + k + ns) % ns;
    const bool wr = k >= 1 && k <= Ls;  // lanes that write
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(nu, FCPT_VISCOSITY);
    FLD(S, FCPT_SIGMA);
    const Fld<SPLIT> vr_o = alt_vr<SPLIT>(P);
    const Fld<SPLIT> vaz_o = alt_vaz<SPLIT>(P);
    const bool iso = !P.adiabatic;  // nu is a ring profile
    const bool store = !ENERGY && P.adiabatic && P.heating_viscous;
    const double dt = P.s->dt;
    // the lane's values of ring r (r < nr); ring nr holds only v_r
    struct Raw {
        double vr, va, s, n;
    };
    auto load = [&](int r) {
        Raw w;
        w.vr = vr(r, j);
        if (r < nr) {
            w.va = vaz(r, j);
            w.s = S(r, j);
            w.n = iso ? cst(P.iso_nu)[r] : nu(r, j);
        } else {
            w.va = w.s = w.n = 0.0;
        }
        return w;
    };
    auto ring = [&](const Raw &w) {
        ViscRing g;
        g.vr = w.vr;
        g.va = w.va;
        g.s = w.s;
        g.n = w.n;
        g.vrL = from_left(w.vr);
        g.sL = from_left(w.s);
        g.nL = iso ? w.n : from_left(w.n);
        g.vaR = from_right(w.va);
        return g;
    };
    // tau_rp at corner (r, j) from rings a = r - 1 and b = r
    auto corner = [&](int r, const ViscRing &a, const ViscRing &b) {
        return stress_rp(P, r, b.va, a.va, b.vr, b.vrL, b.n, a.n, b.nL, a.nL, b.s, a.s, b.sL, a.sL);
    };
    ViscRing A, B;  // rings i, i + 1
    StressDiag dprev{0.0, 0.0, 0.0};
    double trp_c = 0.0, s_prev = 0.0;  // tau_rp(i, j), Sigma(i - 1, j)
    if (i0 >= 1) {
        const ViscRing Z = ring(load(i0 - 1));
        A = ring(load(i0));
        dprev = stress_diag(P, i0 - 1, Z.vr, A.vr, Z.va, Z.vaR, Z.n, Z.s);
        trp_c = corner(i0, Z, A);
        s_prev = Z.s;
    } else {
        A = ring(load(i0));
    }
    B = ring(load(i0 + 1));
    for (int i = i0; i < i1; ++i) {
        const Raw Cn = load(i + 2 <= nr ? i + 2 : nr);  // prefetch, before this ring's stores
        const bool upd_p = i >= 1 && i < nr - 1;
        const bool upd_r = i >= 2 && upd_p;
        const StressDiag d = stress_diag(P, i, A.vr, B.vr, A.va, A.vaR, A.n, A.s);
        const double tppj = from_left(d.tpp);  // diagonal of (i, j - 1)
        const double trp_n = from_right(trp_c);  // corner (i, j + 1)
        const double trp1 = i + 1 <= nr - 1 ? corner(i + 1, A, B) : 0.0;
        const double trp1_n = ENERGY ? from_right(trp1) : 0.0;  // corner (i + 1, j + 1)
        if (wr) {
            if (store) {
                fld<SPLIT>(P, FCPT_DIV_V).set(i, j, d.divv);
                fld<SPLIT>(P, FCPT_TAU_R_R).set(i, j, d.trr);
                fld<SPLIT>(P, FCPT_TAU_PHI_PHI).set(i, j, d.tpp);
                if (i >= 1) fld<SPLIT>(P, FCPT_TAU_R_PHI).set(i, j, trp_c);
            }
            if (upd_p)
                vaz_o.set(i, j, visc_vphi_new<SPLIT>(P, i, dt, A.va, A.s, A.sL, trp_c, trp1, d.tpp, tppj));
            else
                vaz_o.set_raw(i, j, vaz.raw(i, j));
            if (upd_r)
                vr_o.set(i, j, visc_vr_new<SPLIT>(P, i, dt, A.vr, A.s, s_prev, d.trr, dprev.trr, trp_c, trp_n,
                                           d.tpp, dprev.tpp));
            else
                vr_o.set_raw(i, j, vr.raw(i, j));
            if constexpr (ENERGY) {
                const double e = P.e_partner ? buf<SPLIT>(P, P.ealt, P.bg[FCPT_ENERGY])(i, j)
                                             : fld<SPLIT>(P, FCPT_ENERGY)(i, j);
                // H as the kick walk formed it from the same E and Sigma
                const double h = lazy_derived<SPLIT>(P) ? adi_derived<SPLIT>(P, i, e, A.s).h
                                                         : fld<SPLIT>(P, FCPT_SCALE_HEIGHT)(i, j);
                double qplus = 0.0, qminus = 0.0;
                if (upd_p) {
                    qminus = cooling_beta_cell<SPLIT>(P, i, j, P.s->time, e, A.s);
                    if (P.heating_viscous && A.n != 0.0)
                        qplus = viscous_heating_q<SPLIT>(P, A.n, A.s, d.trr, d.tpp, d.divv,
                                                  0.25 * (trp_c + trp1 + trp_n + trp1_n));
                }
                substep3_cell<SPLIT>(P, i, j, e, A.s, h, qplus, qminus);
            }
        }
        dprev = d;
        trp_c = trp1;
        s_prev = A.s;
        A = B;
        B = ring(Cn);
    }
    if (i1 == nr && wr) vr_o.set_raw(nr, j, vr.raw(nr, j));  // the outer v_r row, untouched
}
template <bool SPLIT, bool ENERGY>
__global__ void __launch_bounds__(256) k_visc_walk(KParams P, int tile)
{
    const BlockXY bb = xcd_block();
    b_visc_walk<SPLIT, ENERGY>(P, bb.x, bb.y, tile);
}
// batched form: each disk's tile is its resolved visc_tile
template <bool SPLIT, bool ENERGY>
__global__ void __launch_bounds__(256) kb_visc_walk(const KParams *Ps, const BatchEntry *map)
{
    const BatchEntry e = batch_entry(map);
    const KParams &P = batch_disk(Ps, e.disk);
    b_visc_walk<SPLIT, ENERGY>(P, (int)(e.yx & 0xffffu), (int)(e.yx >> 16), P.visc_tile);
}

// ---------------------------------------------------------------------------
// damping (boundary_conditions/damping.cpp:311-427, 754-776) of one value:
// whether ring i of a vector (v_r) or scalar field lies in the inner or the
// outer damping zone, and the value relaxed towards its reference x0
__device__ __forceinline__ bool damp_in_inner(const KParams &P, bool vector, int i)
{
    const double *radius = vector ? P.Rinf : P.Rmed;
    const int size_radial = vector ? P.nr + 1 : P.nr;
    const double lin = P.damping_inner_limit;
    if (i >= size_radial || !((lin > 1.0) && (radius[0] < P.rmin * lin))) return false;
    return i <= (vector ? P.damp_ids[1] : P.damp_ids[0]);
}
__device__ __forceinline__ bool damp_in_outer(const KParams &P, bool vector, int i)
{
    const double *radius = vector ? P.Rinf : P.Rmed;
    const int size_radial = vector ? P.nr + 1 : P.nr;
    const double lout = P.damping_outer_limit;
    if (i >= size_radial || !((lout < 1.0) && (radius[size_radial - 1] > P.rmax * lout))) return false;
    return i >= (vector ? P.damp_ids[3] : P.damp_ids[2]);
}
__device__ __forceinline__ double damp_apply(const KParams &P, bool vector, bool inner, int i, double dt,
                                             double x, double x0)
{
    const double *radius = vector ? P.Rinf : P.Rmed;
    const double RMIN = P.rmin, RMAX = P.rmax;
    double tau, a;
    if (inner) {
        const double lin = P.damping_inner_limit;
        tau = P.damping_time_factor * 2.0 * M_PI / omega_kepler(P, RMIN);
        a = (radius[i] - RMIN * lin) / (RMIN - RMIN * lin);
    } else {
        const double lout = P.damping_outer_limit;
        tau = P.damping_time_factor * 2.0 * M_PI / omega_kepler(P, P.damping_time_radius_outer);
        a = (radius[i] - RMAX * lout) / (RMAX - RMAX * lout);
    }
    const double exp_factor = exp(-dt * (a * a) / tau);
    return (x - x0) * exp_factor + x0;
}

// damping: block row y maps to an inner-zone ring (y < n_inner) or an
// outer-zone ring
template <bool SPLIT>
__device__ __forceinline__ void b_damping(const KParams &P, int bx, int by)
{
    const int n_inner = (P.damp_ids[0] > P.damp_ids[1] ? P.damp_ids[0] : P.damp_ids[1]) + 1;
    const int ns = P.ns;
    const int j = bx * blockDim.x + threadIdx.x;
    const int y = wave_ring(0, by);
    if (j >= ns) return;
    const double dt = P.s->dt;
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
        if (!(inner ? damp_in_inner(P, vector, i) : damp_in_outer(P, vector, i))) continue;
        const Fld<SPLIT> X = fld<SPLIT>(P, F[q]), X0 = fld<SPLIT>(P, F0[q]);
        X.set(i, j, damp_apply(P, vector, inner, i, dt, X(i, j), X0(i, j)));
    }
}
FCPT_PAIR0(damping, 256)

// ghost rings: zero-gradient Sigma/E, reflecting / zero-gradient / outflow v_r,
// Keplerian v_phi (boundary_conditions.cpp:67-97 and the BC files it dispatches to)
// end of a sim::run iteration: time += dt and the monitor counter
// (simulation.cpp:244, 545-551); nothing else in the kernels that call it
// reads them
__device__ __forceinline__ void advance_clock(const KParams &P, int advance_time)
{
    DevScalars *s = P.s;
    const double time_next_monitor = (s->n_monitor + 1) * P.monitor_timestep;
    s->time = s->time + s->dt;
    if (advance_time == 2 && fabs(time_next_monitor - s->time) < 1e-6 * s->last_dt) s->n_monitor++;
}

// the inner / outer ghost rings of sector j (rows 2 and nr-2 are never
// written here, nr >= 8)
// the ghost values of v_r: rows 0 and 1 from v_r(2), rows nr and nr-1 from
// v_r(nr-2), and the Keplerian v_phi of a ghost ring
__device__ __forceinline__ void ghost_vr_inner(const KParams &P, double v2, double &v0, double &v1)
{
    if (P.bc_inner == FCPT_BC_REFLECTING) {
        v0 = -v2;
        v1 = 0;
    } else if (P.bc_inner == FCPT_BC_ZEROGRADIENT) {
        v0 = v2;
        v1 = v2;
    } else {
        v0 = v1 = v2 > 0.0 ? 0.0 : v2;
    }
}
__device__ __forceinline__ void ghost_vr_outer(const KParams &P, double vn2, double &vn, double &vn1)
{
    if (P.bc_outer == FCPT_BC_REFLECTING) {
        vn = -vn2;
        vn1 = 0;
    } else if (P.bc_outer == FCPT_BC_ZEROGRADIENT) {
        vn = vn2;
        vn1 = vn2;
    } else {
        vn = vn1 = vn2 < 0.0 ? 0.0 : vn2;
    }
}
__device__ __forceinline__ double ghost_vaz(const KParams &P, int r)
{
    const double vk = sqrt(P.G * P.mstar / cst(P.Rmed)[r]);
    return 1.0 * vk - cst(P.Rmed)[r] * P.omega_frame;
}

template <bool SPLIT>
__device__ __forceinline__ void ghost_inner(const KParams &P, int j)
{
    FLD(S, FCPT_SIGMA);
    FLD(E, FCPT_ENERGY);
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    const double s1 = S(1, j), e1 = E(1, j), v2 = vr(2, j);
    S.set(0, j, s1);
    E.set(0, j, e1);
    double v0, v1;
    ghost_vr_inner(P, v2, v0, v1);
    vr.set(0, j, v0);
    vr.set(1, j, v1);
    vaz.set(0, j, ghost_vaz(P, 0));
}

template <bool SPLIT>
__device__ __forceinline__ void ghost_outer(const KParams &P, int j)
{
    const int nr = P.nr;
    FLD(S, FCPT_SIGMA);
    FLD(E, FCPT_ENERGY);
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    const double s2 = S(nr - 2, j), e2 = E(nr - 2, j), vn2 = vr(nr - 2, j);
    S.set(nr - 1, j, s2);
    E.set(nr - 1, j, e2);
    double vn, vn1;
    ghost_vr_outer(P, vn2, vn, vn1);
    vr.set(nr, j, vn);
    vr.set(nr - 1, j, vn1);
    vaz.set(nr - 1, j, ghost_vaz(P, nr - 1));
}

template <bool SPLIT>
__device__ __forceinline__ void b_ghost_bc(const KParams &P, int bx, int by, int advance_time)
{
    const int j = bx * blockDim.x + threadIdx.x;
    if (advance_time && j == 0) advance_clock(P, advance_time);
    if (j >= P.ns) return;
    ghost_inner<SPLIT>(P, j);
    ghost_outer<SPLIT>(P, j);
}
FCPT_PAIR1(ghost_bc, 256, int, advance_time)

// the boundary call before the transport (ghost rings, b_ghost_bc) and the
// transport's ring mean with the FARGO shift (b_ring_mean, mode 1) in one
// launch: the first blocks set the ghost rings (one thread per sector), the
// other blocks (one of 256 threads per ring) form the ring means.  The ghost
// rings' v_phi is one Keplerian value per ring, so their blocks sum that value
// as stored in place of a load (the same sums, the same tree) and no block
// reads what another writes.  Only the azimuthal sweep reads the ring means,
// so they may be formed before the radial sweep.
template <bool SPLIT>
__device__ __forceinline__ void b_ghost_ring_mean(const KParams &P, int bx)
{
    const int ghost_blocks = (P.ns + 255) / 256;
    if (bx < ghost_blocks) {
        const int j = bx * 256 + (int)threadIdx.x;
        if (j < P.ns) {
            ghost_inner<SPLIT>(P, j);
            ghost_outer<SPLIT>(P, j);
        }
        return;
    }
    const int i = bx - ghost_blocks;
    if (i == 0 || i == P.nr - 1) {
        __shared__ double2 sh[256];
        // the ghost ring's v_phi as stored by ghost_inner / ghost_outer in
        // this launch (split storage with stochastic rounding: per sector)
        const Fld<SPLIT> vaz = fld<SPLIT>(P, FCPT_VAZI);
        const double g = ghost_vaz(P, i);
        double acc = 0.0, aacc = 0.0;
        for (int j = (int)threadIdx.x; j < P.ns; j += 256) {
            const double gj = vaz.stored(i, j, g);
            acc += gj;
            aacc += fabs(gj);
        }
        const double2 r = block_sum_abs<256>(acc, aacc, sh);
        ring_mean_store(P, i, r.x, r.y, 1, [&]() {
            double seq = 0.0;
            for (int j = 0; j < P.ns; ++j) seq += vaz.stored(i, j, g);
            return seq;
        });
        return;
    }
    b_ring_mean<SPLIT>(P, i, 0, 1);
}
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_ghost_ring_mean(KParams P)
{
    b_ghost_ring_mean<SPLIT>(P, (int)blockIdx.x);
}
template <bool SPLIT>
__global__ void __launch_bounds__(256) kb_ghost_ring_mean(const KParams *Ps, const BatchEntry *map)
{
    const BatchEntry e = batch_entry(map);
    b_ghost_ring_mean<SPLIT>(batch_disk(Ps, e.disk), (int)(e.yx & 0xffffu));
}

// the final boundary call's ghost rings (b_ghost_bc, after the damping) and
// recalculate_derived_disk_quantities in one pass: a ghost-ring cell first
// takes its neighbour ring's values, then forms its derived fields from what
// it stored, exactly as the two separate kernels would
template <bool SPLIT>
__device__ __forceinline__ void b_final_derived(const KParams &P, int bx, int by, int advance_time)
{
    CELL_LOOP_HEAD(0, P.nr)
    if (advance_time && i == 0 && j == 0) advance_clock(P, advance_time);
    if (i == 0) ghost_inner<SPLIT>(P, j);
    if (i == P.nr - 1) ghost_outer<SPLIT>(P, j);
    derived_store<SPLIT>(P, i, j, derived_load<SPLIT>(P, i, j, true), true, true);
}
FCPT_PAIR1(final_derived, 256, int, advance_time)

// ---------------------------------------------------------------------------
// Transport

// One ring of the radial sweep's window: Sigma, the advected quantities other
// than the density (four momenta, and E when adiabatic: NQ of them), their
// specific values Q / Sigma, the van Leer slopes of Sigma and of each
// specific value, and v_r at the ring's inner interface.  The momenta are
// formed with the expressions of compute_momenta_from_velocities
// (TransportEuler.cpp:471-495).  Rings outside the grid hold Sigma = 1 and
// zero quantities; the reference's interface values never reach them.  The
// density's own specific value Sigma / Sigma is exactly 1 for the positive
// Sigma the floor guarantees, so its interface value is 1 on interior
// interfaces and 0 on the two edge interfaces.
template <int NQ>
struct RingW {
    double s, w[NQ], q[NQ], v;
    double ds, dw[NQ];
};

// The ring's stored values, loaded a ring ahead of their use so that the
// loads are in flight while the previous ring is processed.  Loads go to a
// clamped ring and the results are selected, so the loop body has no branch.
template <bool SPLIT, bool ADI>
struct RingRaw {
    using T = typename Fld<SPLIT>::T;
    T sig, va, vaj, e, v0, v1;
    int r;
};

template <bool SPLIT, bool ADI>
__device__ __forceinline__ void raw_ring(const KParams &P, int r, int j, int jn, RingRaw<SPLIT, ADI> &X)
{
    const bool in = r >= 0 && r < P.nr;
    const int rc = in ? r : 0;
    const int rv = r < 0 ? 0 : (r > P.nr ? P.nr : r);
    const int rv1 = rv < P.nr ? rv + 1 : P.nr;
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    X.r = r;
    X.sig = fld<SPLIT>(P, FCPT_SIGMA).raw(rc, j);
    X.va = vaz.raw(rc, j);
    X.vaj = vaz.raw(rc, jn);
    if constexpr (ADI) X.e = fld<SPLIT>(P, FCPT_ENERGY).raw(rc, j);
    X.v0 = vr.raw(rv, j);
    X.v1 = vr.raw(rv1, j);
}

template <bool SPLIT, bool ADI, int NQ>
__device__ __forceinline__ void decode_ring(const KParams &P, const RingRaw<SPLIT, ADI> &X, RingW<NQ> &R)
{
    const int r = X.r;
    const bool in = r >= 0 && r < P.nr;
    const int rc = in ? r : 0;
    const int rv = r < 0 ? 0 : (r > P.nr ? P.nr : r);
    const int rv1 = rv < P.nr ? rv + 1 : P.nr;
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    const double sig = fld<SPLIT>(P, FCPT_SIGMA).dec(rc, X.sig);
    const double rr = cst(P.Rmed)[rc];
    const double OmegaF = P.omega_frame;
    const double v0 = vr.dec(rv, X.v0);  // v_r(r), r inside the grid or r == nr
    double q[NQ];
    q[0] = sig * vr.dec(rv1, X.v1);      // v_r(r + 1) for r inside the grid
    q[1] = sig * v0;
    q[2] = sig * (vaz.dec(rc, X.vaj) + rr * OmegaF) * rr;
    q[3] = sig * (vaz.dec(rc, X.va) + rr * OmegaF) * rr;
    if constexpr (ADI) q[4] = fld<SPLIT>(P, FCPT_ENERGY).dec(rc, X.e);
    R.v = (r >= 0 && r <= P.nr) ? v0 : 0.0;
    R.s = in ? sig : 1.0;
    const double inv_sig = SPLIT ? recip<true>(sig) : 0.0;
    const SharedDen dsig = shared_den(sig);
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        R.q[k] = in ? q[k] : 0.0;
        const double w = SPLIT ? q[k] * inv_sig : shared_div(q[k], dsig);
        R.w[k] = in ? w : 0.0;
    }
}

// van Leer slopes of ring k from rings k-1, k, k+1 (compute_star_radial first
// loop, TransportEuler.cpp:355-371); zero on the ghost rings
template <bool SPLIT, int NQ>
__device__ __forceinline__ void ring_slopes(const KParams &P, int k, const RingW<NQ> &M, RingW<NQ> &C,
                                            const RingW<NQ> &N)
{
    const bool act = k > 0 && k < P.nr - 1;
    const int kc = act ? k : 1;
    const double im = cst(P.InvDiffRmed)[kc], ip = cst(P.InvDiffRmed)[kc + 1];
    const double ds = vl_lim<SPLIT>((N.s - C.s) * ip, (C.s - M.s) * im);
    C.ds = act ? ds : 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double dw = vl_lim<SPLIT>((N.w[q] - C.w[q]) * ip, (C.w[q] - M.w[q]) * im);
        C.dw[q] = act ? dw : 0.0;
    }
}

// Upwind interface values at interface k between ring A = k-1 and ring B = k
// (compute_star_radial second loop, TransportEuler.cpp:376-405; zero on the
// edge interfaces k = 0 and k = nr) and the flux dt dphi Rinf[k] Q* Sigma* v_r
// of each quantity through it (VanLeerRadial, :545-619).  The reference's
// varq_sup of ring k-1 and varq_inf of ring k are this same product, Rsup[k-1]
// and Rinf[k] being the same radius.
template <int NQ>
__device__ __forceinline__ void interface_flux(const KParams &P, int k, const RingW<NQ> &A,
                                               const RingW<NQ> &B, double dt, double *f, double &fs)
{
    const bool act = k > 0 && k < P.nr;
    const int kc = act ? k : 1;
    const double v = B.v;
    const double fpos = (cst(P.Rmed)[kc] - cst(P.Rmed)[kc - 1] - v * dt) * 0.5;
    const double fneg = (cst(P.Rmed)[kc + 1] - cst(P.Rmed)[kc] + v * dt) * 0.5;
    const bool up = v > 0.0;
    const double ds0 = up ? A.s + fpos * A.ds : B.s - fneg * B.ds;
    const double ds = act ? ds0 : 0.0;
    const double dens = act ? 1.0 : 0.0;
    const double c = dt * P.dphi * cst(P.Rinf)[k];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double qs0 = up ? A.w[q] + fpos * A.dw[q] : B.w[q] - fneg * B.dw[q];
        const double qs = act ? qs0 : 0.0;
        f[q] = c * qs * ds * v;
    }
    fs = c * dens * ds * v;
}

// OneWindRad, TransportEuler.cpp:138-168 and VanLeerRadial :545-619.  Every
// advected quantity of ring i is updated from rings i-2..i+2 of itself and of
// the density before the sweep (DENSITY_INT), so all of them are done in one
// pass that reads the state and writes separate outputs.  A thread owns one
// sector and walks `tile` consecutive rings outward.  Its window holds rings
// k-1 and k with their slopes and ring k+1: each ring is loaded once, each
// slope and each interface flux formed once, and the upwind choice is a
// select, so the loop body is one branch-free block the compiler can schedule
// as a whole.
template <bool SPLIT, bool ADI>
__device__ __forceinline__ void radial_sweep(const KParams &P, int bx, int by, int tile)
{
    constexpr int NQ = ADI ? 5 : 4;
    const int ns = P.ns, nr = P.nr;
    const int j = bx * 64 + (int)threadIdx.x;
    const int i0 = by * tile;
    if (j >= ns || i0 >= nr) return;
    const int jn = j == ns - 1 ? 0 : j + 1;
    const double dt = P.s->dt;
    RingW<NQ> A, B, C, D;
    {
        RingRaw<SPLIT, ADI> X0, X1, X2, X3;
        raw_ring<SPLIT, ADI>(P, i0 - 2, j, jn, X0);
        raw_ring<SPLIT, ADI>(P, i0 - 1, j, jn, X1);
        raw_ring<SPLIT, ADI>(P, i0, j, jn, X2);
        raw_ring<SPLIT, ADI>(P, i0 + 1, j, jn, X3);
        decode_ring<SPLIT, ADI, NQ>(P, X0, D);
        decode_ring<SPLIT, ADI, NQ>(P, X1, A);
        decode_ring<SPLIT, ADI, NQ>(P, X2, B);
        decode_ring<SPLIT, ADI, NQ>(P, X3, C);
    }
    ring_slopes<SPLIT, NQ>(P, i0 - 1, D, A, B);
    ring_slopes<SPLIT, NQ>(P, i0, A, B, C);
    double f_prev[NQ], fs_prev;
    interface_flux<NQ>(P, i0, A, B, dt, f_prev, fs_prev);
    const int iend = i0 + tile < nr ? i0 + tile : nr;
    RingRaw<SPLIT, ADI> X;
    raw_ring<SPLIT, ADI>(P, i0 + 2, j, jn, X);
    for (int k = i0 + 1; k <= iend; ++k) {
        decode_ring<SPLIT, ADI, NQ>(P, X, D);      // ring k+1
        raw_ring<SPLIT, ADI>(P, k + 2, j, jn, X);  // in flight during this iteration
        ring_slopes<SPLIT, NQ>(P, k, B, C, D);            // ring k
        A = B;
        B = C;
        C = D;
        double f[NQ], fs;
        interface_flux<NQ>(P, k, A, B, dt, f, fs);
        // ring k-1 lies between interfaces k-1 and k
        const int i = k - 1;
        const double invs = cst(P.InvSurf)[i];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            buf<SPLIT>(P, P.mom_b[q], P.bgm[q], kBufMom + q).set(i, j, A.q[q] + (f_prev[q] - f[q]) * invs);
        if constexpr (ADI) buf<false>(P, P.en_b, nullptr).set(i, j, A.q[4] + (f_prev[4] - f[4]) * invs);
        buf<false>(P, P.sig_b, nullptr).set(i, j, A.s + (fs_prev - fs) * invs);
#pragma unroll
        for (int q = 0; q < NQ; ++q) f_prev[q] = f[q];
        fs_prev = fs;
    }
}

// the radial sweep under two register budgets: 3 wavefronts per SIMD (168
// VGPRs) or 2 (256 VGPRs)
template <bool SPLIT, bool ADI>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_radial_sweep2(KParams P, int tile)
{
    const BlockXY bb = xcd_block();
    radial_sweep<SPLIT, ADI>(P, bb.x, bb.y, tile);
}
template <bool SPLIT, bool ADI>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) k_radial_sweep3(KParams P, int tile)
{
    const BlockXY bb = xcd_block();
    radial_sweep<SPLIT, ADI>(P, bb.x, bb.y, tile);
}
// batched forms: each disk's tile is its resolved radial_tile
template <bool SPLIT, bool ADI>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
    kb_radial_sweep2(const KParams *Ps, const BatchEntry *map)
{
    const BatchEntry e = batch_entry(map);
    const KParams &P = batch_disk(Ps, e.disk);
    radial_sweep<SPLIT, ADI>(P, (int)(e.yx & 0xffffu), (int)(e.yx >> 16), P.radial_tile);
}
template <bool SPLIT, bool ADI>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
    kb_radial_sweep3(const KParams *Ps, const BatchEntry *map)
{
    const BatchEntry e = batch_entry(map);
    const KParams &P = batch_disk(Ps, e.disk);
    radial_sweep<SPLIT, ADI>(P, (int)(e.yx & 0xffffu), (int)(e.yx >> 16), P.radial_tile);
}

// van Leer slope of a sector from its two neighbours (the slope half of
// ComputeStarTheta, TransportEuler.cpp:416-465)
template <bool SPLIT>
__device__ __forceinline__ double slope_theta(double q, double invdxtheta)
{
    return 0.5 * vl_lim<SPLIT>(from_right(q) - q, q - from_left(q)) * invdxtheta;
}

// upwind interface value between sector j-1 and sector j (the second half of
// ComputeStarTheta): q, dq of sector j in this lane, of sector j-1 in the lane
// to the left
__device__ __forceinline__ double star_theta(double q, double dq, double dxtheta, double ksi)
{
    const double lo = from_left(q) + (dxtheta - ksi) * from_left(dq);
    const double hi = q - (dxtheta + ksi) * dq;
    return (ksi > 0.0) ? lo : hi;
}

// OneWindTheta, TransportEuler.cpp:270-289, in one pass per ring segment:
// QuantitiesAdvection (:292-304, VanLeerTheta :630-664) with the residual
// velocity v - <v>, then with the uniform residual of the ring, then
// AdvectSHIFT (:238-268).  Each wavefront owns kAzSegW consecutive sectors of
// one ring, two adjacent sectors per lane, and loads them with kAzHalo
// sectors on either side; a sector's neighbours are the lane's other sector
// or the adjacent lane's (one DPP move per neighbour pair), so the kernel
// needs neither LDS nor barriers.  Each stencil step spoils one more sector
// at either edge; after the two van Leer passes sectors kAzHalo .. 127 -
// kAzHalo of the wavefront hold exact results, and the result of sector j is
// written to sector j + Nshift.  The density's own specific value is exactly
// 1 (positive Sigma), so its interface values are 1 and drop out of its
// flux.  Every slope is formed once per sector and each interface selects its
// upwind side.
constexpr int kAzHalo = 4;
constexpr int kAzSegW = 128 - 2 * kAzHalo;  // sectors written per wavefront
constexpr int kAzWaves = 2;                 // wavefronts per block

// the two sectors of a lane: a = 2 lane, b = 2 lane + 1 (within the segment)
struct D2 {
    double a, b;
};
__device__ __forceinline__ D2 left2(const D2 &x) { return D2{from_left(x.b), x.a}; }
__device__ __forceinline__ D2 right2(const D2 &x) { return D2{x.b, from_right(x.a)}; }

template <bool SPLIT>
__device__ __forceinline__ D2 slope2(const D2 &q, double invdxtheta)
{
    const D2 r = right2(q), l = left2(q);
    return D2{0.5 * vl_lim<SPLIT>(r.a - q.a, q.a - l.a) * invdxtheta,
              0.5 * vl_lim<SPLIT>(r.b - q.b, q.b - l.b) * invdxtheta};
}

// upwind interface value between each sector and its left neighbour (the
// second half of ComputeStarTheta)
__device__ __forceinline__ D2 star2(const D2 &q, const D2 &dq, double dxtheta, const D2 &ksi)
{
    const D2 lq = left2(q), ldq = left2(dq);
    const double loa = lq.a + (dxtheta - ksi.a) * ldq.a, hia = q.a - (dxtheta + ksi.a) * dq.a;
    const double lob = lq.b + (dxtheta - ksi.b) * ldq.b, hib = q.b - (dxtheta + ksi.b) * dq.b;
    return D2{(ksi.a > 0.0) ? loa : hia, (ksi.b > 0.0) ? lob : hib};
}

// star2 for a velocity of one sign on the whole ring (the uniform residual):
// UP (ksi > 0) takes the left sector's side for both sectors, else the own
__device__ __forceinline__ D2 star2_up(const D2 &q, const D2 &dq, double dxtheta, double ksi)
{
    const D2 lq = left2(q), ldq = left2(dq);
    return D2{lq.a + (dxtheta - ksi) * ldq.a, lq.b + (dxtheta - ksi) * ldq.b};
}
__device__ __forceinline__ D2 star2_down(const D2 &q, const D2 &dq, double dxtheta, double ksi)
{
    return D2{q.a - (dxtheta + ksi) * dq.a, q.b - (dxtheta + ksi) * dq.b};
}
template <bool UP>
__device__ __forceinline__ D2 star2_uniform(const D2 &q, const D2 &dq, double dxtheta, double ksi)
{
    return UP ? star2_up(q, dq, dxtheta, ksi) : star2_down(q, dq, dxtheta, ksi);
}

// flux into the sector through its left interface minus flux out through the
// right one
__device__ __forceinline__ D2 flux_div2(const D2 &f)
{
    const D2 r = right2(f);
    double va = f.a, vb = f.b;
    va -= r.a;
    vb -= r.b;
    return D2{va, vb};
}

// Both FARGO passes for this lane's two sectors of ring i (sources at
// sectors jsa, jsb = jsa + 1): put(-1, .) receives the transported density
// and put(q, .) each quantity's transported momentum (q = 0 .. 3) and energy
// (q = 4), before the cyclic shift, as soon as it is formed (so its
// registers are free for the next quantity).  The second pass's velocity is
// the ring's uniform residual, so its upwind side is the same for every
// sector: UP is its sign (ksi > 0), chosen once per ring by the caller.
template <bool SPLIT, bool ADI, bool UP, class Put>
__device__ __forceinline__ void az_ring(const KParams &P, int i, int jsa, int jsb, double dt, Put &&put)
{
    constexpr int NQ = ADI ? 5 : 4;
    const double vm = cst(P.vmean)[i];
    const double vu = cst(P.vuni)[i];
    const double ksu = vu * dt;
    const double dxtheta = P.dphi * cst(P.Rmed)[i];
    const double invdxtheta = 1.0 / dxtheta;
    const double dxrad = (cst(P.Rsup)[i] - cst(P.Rinf)[i]) * dt;
    const double invsurf = cst(P.InvSurf)[i];
    // every global load first
    FLD(vaz, FCPT_VAZI);
    const Fld<false> Sb = buf<false>(P, P.sig_b, nullptr);
    const D2 s0{Sb(i, jsa), Sb(i, jsb)};
    const D2 v0{vaz(i, jsa) - vm, vaz(i, jsb) - vm};
    D2 q0[NQ];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const Fld<SPLIT> m = buf<SPLIT>(P, P.mom_b[q], P.bgm[q], kBufMom + q);
        q0[q] = D2{m(i, jsa), m(i, jsb)};
    }
    if constexpr (ADI) {
        const Fld<false> Eb = buf<false>(P, P.en_b, nullptr);
        q0[4] = D2{Eb(i, jsa), Eb(i, jsb)};
    }
    const D2 ks0{v0.a * dt, v0.b * dt};
    // density: first pass (residual velocity), then second pass (uniform residual)
    const D2 ds0 = star2(s0, slope2<SPLIT>(s0, invdxtheta), dxtheta, ks0);
    const D2 vd0 = flux_div2(D2{dxrad * ds0.a * v0.a, dxrad * ds0.b * v0.b});
    const D2 s1{s0.a + vd0.a * invsurf, s0.b + vd0.b * invsurf};
    const D2 ds1 = star2_uniform<UP>(s1, slope2<SPLIT>(s1, invdxtheta), dxtheta, ksu);
    const D2 vd1 = flux_div2(D2{dxrad * ds1.a * vu, dxrad * ds1.b * vu});
    put(-1, D2{s1.a + vd1.a * invsurf, s1.b + vd1.b * invsurf});
    // the other quantities
    const D2 inv_s0{SPLIT ? recip<true>(s0.a) : 0.0, SPLIT ? recip<true>(s0.b) : 0.0};
    const D2 inv_s1{SPLIT ? recip<true>(s1.a) : 0.0, SPLIT ? recip<true>(s1.b) : 0.0};
    const SharedDen s0a = shared_den(s0.a), s0b = shared_den(s0.b);
    const SharedDen s1a = shared_den(s1.a), s1b = shared_den(s1.b);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const D2 w0{SPLIT ? q0[q].a * inv_s0.a : shared_div(q0[q].a, s0a),
                    SPLIT ? q0[q].b * inv_s0.b : shared_div(q0[q].b, s0b)};
        const D2 qs0 = star2(w0, slope2<SPLIT>(w0, invdxtheta), dxtheta, ks0);
        const D2 vq0 = flux_div2(D2{dxrad * qs0.a * ds0.a * v0.a, dxrad * qs0.b * ds0.b * v0.b});
        const D2 q1{q0[q].a + vq0.a * invsurf, q0[q].b + vq0.b * invsurf};
        const D2 w1{SPLIT ? q1.a * inv_s1.a : shared_div(q1.a, s1a), SPLIT ? q1.b * inv_s1.b : shared_div(q1.b, s1b)};
        const D2 qs1 = star2_uniform<UP>(w1, slope2<SPLIT>(w1, invdxtheta), dxtheta, ksu);
        const D2 vq1 = flux_div2(D2{dxrad * qs1.a * ds1.a * vu, dxrad * qs1.b * ds1.b * vu});
        put(q, D2{q1.a + vq1.a * invsurf, q1.b + vq1.b * invsurf});
    }
}

template <bool SPLIT, bool ADI>
__device__ __forceinline__ void b_azimuthal_sweep(const KParams &P, int bx, int by)
{
    constexpr int NQ = ADI ? 5 : 4;
    const int ns = P.ns;
    const int i = by;
    const int lane = threadIdx.x & 63;
    const int j0 = __builtin_amdgcn_readfirstlane(bx * kAzWaves + (int)(threadIdx.x >> 6)) * kAzSegW;
    if (j0 >= ns) return;  // the whole wavefront
    const int Ls = ns - j0 < kAzSegW ? ns - j0 : kAzSegW;
    const double dt = P.s->dt;
    const int nsh = cst(P.nshift)[i];
    // this lane's sectors (offset k = 2 lane, 2 lane + 1 from the segment's
    // first halo sector) and where their results go
    const int ka = 2 * lane, kb = 2 * lane + 1;
    int jsa = (j0 - kAzHalo + ka) % ns;
    if (jsa < 0) jsa += ns;
    const int jsb = jsa + 1 == ns ? 0 : jsa + 1;
    int jda = (j0 - kAzHalo + ka + nsh) % ns;
    if (jda < 0) jda += ns;
    const int jdb = jda + 1 == ns ? 0 : jda + 1;
    const bool wa = ka >= kAzHalo && ka < kAzHalo + Ls;
    const bool wb = kb >= kAzHalo && kb < kAzHalo + Ls;
    auto put = [&](int q, const D2 &v) {
        if (q < 0) {
            const Fld<false> Sc = buf<false>(P, P.sig_c, nullptr);
            if (wa) Sc.set(i, jda, v.a);
            if (wb) Sc.set(i, jdb, v.b);
        } else if (q < 4) {
            const Fld<false> m = mom_out(P, q);
            if (wa) m.set(i, jda, v.a);
            if (wb) m.set(i, jdb, v.b);
        } else {
            const Fld<false> Ec = buf<false>(P, P.en_c, nullptr);
            if (wa) Ec.set(i, jda, v.a);
            if (wb) Ec.set(i, jdb, v.b);
        }
    };
    if (cst(P.vuni)[i] * dt > 0.0) az_ring<SPLIT, ADI, true>(P, i, jsa, jsb, dt, put);
    else az_ring<SPLIT, ADI, false>(P, i, jsa, jsb, dt, put);
}
template <bool SPLIT, bool ADI>
__global__ void __launch_bounds__(64 * kAzWaves) k_azimuthal_sweep(KParams P)
{
    const BlockXY bb = xcd_block();
    b_azimuthal_sweep<SPLIT, ADI>(P, bb.x, bb.y);
}
template <bool SPLIT, bool ADI>
__global__ void __launch_bounds__(64 * kAzWaves) kb_azimuthal_sweep(const KParams *Ps, const BatchEntry *map)
{
    const BatchEntry e = batch_entry(map);
    b_azimuthal_sweep<SPLIT, ADI>(batch_disk(Ps, e.disk), (int)(e.yx & 0xffffu), (int)(e.yx >> 16));
}

// compute_velocities_from_momenta (TransportEuler.cpp:498-535), then
// assure_minimum_value(Sigma) (SourceEuler.cpp:102-131) and, adiabatic,
// SetTemperatureFloorCeilValues (TransportEuler.cpp:127-131)
template <bool SPLIT>
__device__ __forceinline__ void b_velocities(const KParams &P, int bx, int by)
{
    CELL_LOOP_HEAD(0, P.nr)
    const Fld<false> Sb = buf<false>(P, P.sig_c, nullptr);
    const Fld<false> m0 = mom_out(P, 0);
    const Fld<false> m1 = mom_out(P, 1);
    const Fld<false> m2 = mom_out(P, 2);
    const Fld<false> m3 = mom_out(P, 3);
    const Fld<false> Eb = buf<false>(P, P.en_c, nullptr);
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    const int jm = j == 0 ? ns - 1 : j - 1;
    const int im = i >= 1 ? i - 1 : 0;
    const double s0 = Sb(i, j), sm = Sb(im, j), sjm = Sb(i, jm);
    const double m0v = m0(im, j), m1v = m1(i, j), m2v = m2(i, jm), m3v = m3(i, j);
    const double e = P.adiabatic ? Eb(i, j) : 0.0;
    if (i == 0) {
        vr.set(i, j, 0.0);
    } else {
        vr.set(i, j, (m0v + m1v) / (sm + s0));
    }
    vaz.set(i, j, (m2v + m3v) / (sjm + s0) * cst(P.InvRmed)[i] - cst(P.Rmed)[i] * P.omega_frame);
    const double floor_value = P.sigma_floor * P.sigma0;
    double sig = s0;
    if (sig < floor_value) sig = floor_value;
    fld<SPLIT>(P, FCPT_SIGMA).set(i, j, sig);
    if (P.adiabatic) fld<SPLIT>(P, FCPT_ENERGY).set(i, j, temperature_floor<SPLIT>(P, e, sig));
}
FCPT_PAIR0(velocities, 256)

// the damping of the final boundary call applied to a value the velocity
// kernel stores: the value is damped as a later load of it returns it, so
// the field is that of the velocity kernel followed by the damping kernel
template <bool SPLIT>
__device__ __forceinline__ double damp_stored(const KParams &P, const Fld<SPLIT> &X, int f0, bool vector,
                                              int i, int j, double dt, double x)
{
    if (!P.damping) return x;
    const bool in = damp_in_inner(P, vector, i), out = damp_in_outer(P, vector, i);
    if (!in && !out) return x;
    x = X.stored(i, j, x);
    const double x0 = fld<SPLIT>(P, f0)(i, j);
    if (in) x = damp_apply(P, vector, true, i, dt, x, x0);
    if (out) x = damp_apply(P, vector, false, i, dt, x, x0);
    return x;
}

// The end of the fused Euler step in one launch: the velocity kernel, the
// final boundary call (damping, ghost rings, the clock) and
// recalculate_derived_disk_quantities.  A cell forms its velocities, Sigma
// and E as the velocity kernel and the damping store them, and its derived
// fields from the stored values, as k_final_derived reads them.  The ghost
// rings copy their neighbour ring, which another thread stores, so a
// ghost-ring thread forms the neighbour's values itself (the same
// expressions on the same inputs):
// ring 0 takes Sigma and E of ring 1 and v_r from v_r(2), ring nr-1 those of
// ring nr-2; v_r(1) is a ghost value too, so the ring-1 thread stores that
// one instead of its own, and no cell is stored by two threads.
template <bool SPLIT>
__device__ __forceinline__ double vel_final_vr(const KParams &P, int r, int j, double dt)
{
    // v_r(r >= 1) as the velocity kernel and the damping pass it to set()
    const Fld<false> Sb = buf<false>(P, P.sig_c, nullptr);
    const Fld<false> m0 = mom_out(P, 0);
    const Fld<false> m1 = mom_out(P, 1);
    FLD(vr, FCPT_VRAD);
    const double den = Sb(r - 1, j) + Sb(r, j);
    const double v = SPLIT ? (m0(r - 1, j) + m1(r, j)) * recip<true>(den) : (m0(r - 1, j) + m1(r, j)) / den;
    return damp_stored<SPLIT>(P, vr, FCPT_VRAD0, true, r, j, dt, v);
}
template <bool SPLIT>
__device__ __forceinline__ void vel_final_sig_e(const KParams &P, int r, int j, double dt, double &sig,
                                                double &e, bool need_e)
{
    // Sigma and (adiabatic) E of ring r as the velocity kernel and the
    // damping pass them to set(); isothermal E is not touched there and is
    // the stored value
    FLD(S, FCPT_SIGMA);
    FLD(E, FCPT_ENERGY);
    const double floor_value = P.sigma_floor * P.sigma0;
    double s0 = buf<false>(P, P.sig_c, nullptr)(r, j);
    if (s0 < floor_value) s0 = floor_value;
    sig = damp_stored<SPLIT>(P, S, FCPT_SIGMA0, false, r, j, dt, s0);
    if (P.adiabatic) {
        const double eb = buf<false>(P, P.en_c, nullptr)(r, j);
        e = damp_stored<SPLIT>(P, E, FCPT_ENERGY0, false, r, j, dt, temperature_floor<SPLIT>(P, eb, s0));
    } else {
        e = need_e ? E(r, j) : 0.0;
    }
}

template <bool SPLIT>
__device__ __forceinline__ void b_velocities_final(const KParams &P, int bx, int by, int advance_time)
{
    CELL_LOOP_HEAD(0, P.nr)
    const int nr = P.nr;
    const double dt = P.s->dt;
    FLD(vr, FCPT_VRAD);
    FLD(vaz, FCPT_VAZI);
    FLD(S, FCPT_SIGMA);
    FLD(E, FCPT_ENERGY);
    const bool ghost = i == 0 || i == nr - 1;
    // every value first, every store after (a load after a store waits for it)
    double sr, er, va = 0.0, vb = 0.0, vphi = 0.0;
    if (ghost) {
        const int r = i == 0 ? 1 : nr - 2;  // the ring the ghost ring copies
        vel_final_sig_e<SPLIT>(P, r, j, dt, sr, er, true);
        // the ghost ring takes the neighbour's values as a load returns them
        sr = S.stored(r, j, sr);
        if (P.adiabatic) er = E.stored(r, j, er);
        if (i == 0) ghost_vr_inner(P, vr.stored(2, j, vel_final_vr<SPLIT>(P, 2, j, dt)), va, vb);
        else ghost_vr_outer(P, vr.stored(nr - 2, j, vel_final_vr<SPLIT>(P, nr - 2, j, dt)), va, vb);
        vphi = ghost_vaz(P, i);
    } else {
        const Fld<false> Sb = buf<false>(P, P.sig_c, nullptr);
        const Fld<false> m2 = mom_out(P, 2);
        const Fld<false> m3 = mom_out(P, 3);
        const int jm = j == 0 ? ns - 1 : j - 1;
        const double s0 = Sb(i, j), sjm = Sb(i, jm);
        const double msum = m2(i, jm) + m3(i, j);
        const double v_phi = (SPLIT ? msum * recip<true>(sjm + s0) : msum / (sjm + s0)) * cst(P.InvRmed)[i] -
                             cst(P.Rmed)[i] * P.omega_frame;
        vphi = damp_stored<SPLIT>(P, vaz, FCPT_VAZI0, false, i, j, dt, v_phi);
        if (i == 1) ghost_vr_inner(P, vr.stored(2, j, vel_final_vr<SPLIT>(P, 2, j, dt)), va, vb);
        else vb = vel_final_vr<SPLIT>(P, i, j, dt);
        vel_final_sig_e<SPLIT>(P, i, j, dt, sr, er, false);
    }
    // stores: the ghost rings, v_r(1) of the ring-1 thread, or the cell's own
    const double s_stored = S.set_get(i, j, sr);
    double e_stored = 0.0;
    if (ghost || P.adiabatic) e_stored = E.set_get(i, j, er);
    vaz.set(i, j, vphi);
    if (i == 0) {
        vr.set(0, j, va);
    } else if (i == nr - 1) {
        vr.set(nr, j, va);
        vr.set(nr - 1, j, vb);
    } else {
        vr.set(i, j, vb);  // v_r(1): the ghost value
    }
    if (advance_time && i == 0 && j == 0) advance_clock(P, advance_time);
    DerivedIn d{P.adiabatic ? e_stored : 0.0, s_stored, 0.0, 0.0};
    if (!P.adiabatic) {
        d.cs = cst(P.iso_cs)[i];
        d.h = cst(P.iso_h)[i];
    }
    derived_store<SPLIT>(P, i, j, d, true, true);
}
FCPT_PAIR1(velocities_final, 256, int, advance_time)

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
    const bool active = i >= 1 && i < P.nr - 1 && cst(P.Rmed)[i] <= radius;
    if (active) {
        FLD(S, FCPT_SIGMA);
        FLD(vaz, FCPT_VAZI);
        for (int j = t; j < ns; j += 64) {
            const int jm = j == 0 ? ns - 1 : j - 1;
            const double s = S(i, j);
            neumaier_add(m, mc, cst(P.Surf)[i] * s);
            neumaier_add(l, lc, cst(P.Surf)[i] * 0.5 * (s + S(i, jm)) * cst(P.Rmed)[i] *
                                    (vaz(i, j) + P.omega_frame * cst(P.Rmed)[i]));
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

// streaming copy of four work arrays (bandwidth probe)
__global__ void __launch_bounds__(256) k_copy4(KParams P, size_t n8)
{
    // one 8-byte element of each of the four arrays per thread, every load
    // before the stores: the load and store forms of the step kernels (on
    // MI355X this form copies ~25 % faster than a grid-stride loop of 16-byte
    // elements)
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n8) return;
    double v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = static_cast<const double *>(P.mom[q])[k];
#pragma unroll
    for (int q = 0; q < 4; ++q) static_cast<double *>(P.mom_b[q])[k] = v[q];
}

// split storage: re-express a field's deviations about a new ring background
__global__ void __launch_bounds__(256) k_rebase(float *d, const double *oldbg, const double *newbg,
                                                int ns, int rows)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = __builtin_amdgcn_readfirstlane((int)(blockIdx.y * blockDim.y + threadIdx.y));
    if (j >= ns || i >= rows) return;
    const size_t c = IX(i, j);
    d[c] = (float)((oldbg[i] + (double)d[c]) - newbg[i]);
}

// split storage: the ring means of a field, background + mean deviation (one
// block per row; the deviations summed in FP64, pairwise tree order)
__global__ void __launch_bounds__(256) k_split_ring_mean(const float *d, const double *bg, int ns,
                                                         double *out)
{
    __shared__ double sh[256];
    const int i = blockIdx.x;
    const float *row = d + (size_t)i * (size_t)ns;
    double acc = 0.0;
    strided_row<256>(ns, (int)threadIdx.x, [&](int j) { return (double)row[j]; }, [&](double x) { acc += x; });
    const double sum = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) out[i] = bg[i] + sum / ns;
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
size_t field_element_size(const KParams &P, int field)
{
    return field == FCPT_POTENTIAL ? sizeof(double) : element_size(P);
}

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

void init_launch_fields(KParams &P)
{
    const dim3 g = cfl_grid(P);
    P.cfl_gx = (int)g.x;
    P.cfl_gy = (int)g.y;
}

void launch_cfl(const KParams &P, hipStream_t st, int mode)
{
    LAUNCH(k_cfl_rings, dim3((unsigned)P.nr), dim3(512), st, P);
    hipLaunchKernelGGL(k_cfl_finalize_rings, dim3(1), dim3(256), 0, st, P, mode);
}

void launch_cfl_cells(const KParams &P, hipStream_t st)
{
    LAUNCH(k_cfl_rings, dim3((unsigned)P.nr), dim3(512), st, P);
}

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
void launch_set_params(const KParams &P, hipStream_t st, const fcpt_step_params &sp,
                       const BodySmoothing &sm)
{
    hipLaunchKernelGGL(k_set_params, dim3(1), dim3(1), 0, st, P, sp, sm);
}
void launch_set_mid_params(const KParams &P, hipStream_t st, const fcpt_step_params &sp,
                           const BodySmoothing &sm)
{
    hipLaunchKernelGGL(k_set_mid_params, dim3(1), dim3(1), 0, st, P, sp, sm);
}
void launch_leapfrog_phase(const KParams &P, hipStream_t st, int phase)
{
    hipLaunchKernelGGL(k_leapfrog_phase, dim3(1), dim3(1), 0, st, P, phase);
}
void launch_pressure(const KParams &P, hipStream_t st) { LAUNCH(k_pressure, cell_grid(P, P.nr), kBlock, st, P); }

// a 2-D field whose rows are constant: row i = ring[i] (split storage: the
// deviation from bg[i])
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_fill_rows(void *dst, const double *ring, const double *bg,
                                                   int ns, int rows)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = wave_ring(0, (int)blockIdx.y);
    if (j >= ns || i >= rows) return;
    const size_t c = (size_t)i * (size_t)ns + (size_t)j;
    if constexpr (SPLIT) static_cast<float *>(dst)[c] = (float)(ring[i] - bg[i]);
    else static_cast<double *>(dst)[c] = ring[i];
}

void launch_fill_rows(const KParams &P, hipStream_t st, void *dst, const double *ring,
                      const double *bg, int rows)
{
    const dim3 g((unsigned)((P.ns + 63) / 64), (unsigned)((rows + 3) / 4));
    if (P.split) hipLaunchKernelGGL(k_fill_rows<true>, g, kBlock, 0, st, dst, ring, bg, P.ns, rows);
    else hipLaunchKernelGGL(k_fill_rows<false>, g, kBlock, 0, st, dst, ring, bg, P.ns, rows);
}
void launch_copy_velocities(const KParams &P, hipStream_t st)
{
    LAUNCH(k_copy_velocities, cell_grid(P, P.nr + 1), kBlock, st, P);
}

void launch_potential(const KParams &P, hipStream_t st) { launch_potential_fold(P, st, -1); }

void launch_potential_fold(const KParams &P, hipStream_t st, int fold_mode)
{
    const dim3 g = cell_grid(P, P.nr);
    const unsigned gx = (unsigned)((P.ns + 64 * kPotCells - 1) / (64 * kPotCells));
    LAUNCH(k_potential_fold, dim3(gx, g.y + 1), kBlock, st, P, fold_mode);
}

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
        LAUNCH(k_damping, cell_grid(P, n_inner + n_outer), kBlock, st, P);
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
    // register budget: 3 wavefronts per SIMD (168 VGPRs) or 2 (256 VGPRs)
    const bool w2 = P.radial_waves == 2;
    if (P.split) {
        if (P.adiabatic) {
            if (w2) hipLaunchKernelGGL((k_radial_sweep2<true, true>), g, dim3(64), 0, st, P, tile);
            else hipLaunchKernelGGL((k_radial_sweep3<true, true>), g, dim3(64), 0, st, P, tile);
        } else {
            if (w2) hipLaunchKernelGGL((k_radial_sweep2<true, false>), g, dim3(64), 0, st, P, tile);
            else hipLaunchKernelGGL((k_radial_sweep3<true, false>), g, dim3(64), 0, st, P, tile);
        }
    } else {
        if (P.adiabatic) {
            if (w2) hipLaunchKernelGGL((k_radial_sweep2<false, true>), g, dim3(64), 0, st, P, tile);
            else hipLaunchKernelGGL((k_radial_sweep3<false, true>), g, dim3(64), 0, st, P, tile);
        } else {
            if (w2) hipLaunchKernelGGL((k_radial_sweep2<false, false>), g, dim3(64), 0, st, P, tile);
            else hipLaunchKernelGGL((k_radial_sweep3<false, false>), g, dim3(64), 0, st, P, tile);
        }
    }
}
void launch_azimuthal_sweep(const KParams &P, hipStream_t st)
{
    const int waves = (P.ns + kAzSegW - 1) / kAzSegW;
    const dim3 g((unsigned)((waves + kAzWaves - 1) / kAzWaves), (unsigned)P.nr);
    const dim3 b(64 * kAzWaves);
    if (P.split) {
        if (P.adiabatic) hipLaunchKernelGGL((k_azimuthal_sweep<true, true>), g, b, 0, st, P);
        else hipLaunchKernelGGL((k_azimuthal_sweep<true, false>), g, b, 0, st, P);
    } else {
        if (P.adiabatic) hipLaunchKernelGGL((k_azimuthal_sweep<false, true>), g, b, 0, st, P);
        else hipLaunchKernelGGL((k_azimuthal_sweep<false, false>), g, b, 0, st, P);
    }
}

void launch_copy4(const KParams &P, hipStream_t st)
{
    // the bytes of four fields, as doubles
    const size_t n8 = (size_t)P.nr * P.ns * element_size(P) / 8;
    const size_t blocks = (n8 + 255) / 256;
    hipLaunchKernelGGL(k_copy4, dim3((unsigned)(blocks > 0 ? blocks : 1)), dim3(256), 0, st, P, n8);
}

void launch_transport(const KParams &P, hipStream_t st) { launch_transport_to(P, P, st); }

// the sweeps read the velocities of view P; the velocity kernel reads only the
// transported density and momenta and writes the velocities of view out
void launch_transport_to(const KParams &P, const KParams &out, hipStream_t st)
{
    launch_radial_sweep(P, st);
    launch_ring_mean(P, st, 1);
    launch_azimuthal_sweep(P, st);
    LAUNCH(k_velocities, cell_grid(out, out.nr), kBlock, st, out);
}

void launch_transport_ghost(const KParams &P, const KParams &out, hipStream_t st, int final_advance)
{
    const int gb = (P.ns + 255) / 256;
    LAUNCH(k_ghost_ring_mean, dim3((unsigned)(gb + P.nr)), dim3(256), st, P);
    launch_radial_sweep(P, st);
    launch_azimuthal_sweep(P, st);
    if (final_advance) LAUNCH(k_velocities_final, cell_grid(out, out.nr), kBlock, st, out, final_advance);
    else LAUNCH(k_velocities, cell_grid(out, out.nr), kBlock, st, out);
}

KParams swapped_view(const KParams &P)
{
    KParams Q = P;
    Q.f[FCPT_VRAD] = P.valt[0];
    Q.valt[0] = P.f[FCPT_VRAD];
    Q.f[FCPT_VAZI] = P.valt[1];
    Q.valt[1] = P.f[FCPT_VAZI];
    return Q;
}

void launch_src_fused(const KParams &P, hipStream_t st)
{
    LAUNCH(k_src_fused, cell_grid(P, P.nr + 1), kBlock, st, P);
}
void launch_src_energy(const KParams &P, hipStream_t st)
{
    LAUNCH(k_src_energy, cell_grid(P, P.nr - 1), kBlock, st, P);
}
void launch_av_fused(const KParams &P, hipStream_t st, int with_viscosity)
{
    LAUNCH(k_av_fused, cell_grid(P, P.nr + 1), kBlock, st, P, with_viscosity);
}
// rings per thread of the viscous walk: 8, or 4, while that gives 3072
// wavefronts (the window start costs two extra rings per walk), else 1 (the
// cell-parallel form)
static int visc_tile(const KParams &P)
{
    if (P.visc_tile > 0) return P.visc_tile;
    const long segs = (P.ns + kViscSegW - 1) / kViscSegW;
    int tile = 8;
    while (tile > 1 && segs * ((P.nr + tile - 1) / tile) < 3072) tile = tile > 4 ? tile / 2 : 1;
    return tile;
}

static void launch_visc_walk(const KParams &P, int tile, bool energy, hipStream_t st)
{
    const int tiles = (P.nr + tile - 1) / tile;
    const dim3 g((unsigned)((P.ns + kViscSegW - 1) / kViscSegW), (unsigned)((tiles + 3) / 4));
    if (P.split) {
        if (energy) hipLaunchKernelGGL((k_visc_walk<true, true>), g, kBlock, 0, st, P, tile);
        else hipLaunchKernelGGL((k_visc_walk<true, false>), g, kBlock, 0, st, P, tile);
    } else {
        if (energy) hipLaunchKernelGGL((k_visc_walk<false, true>), g, kBlock, 0, st, P, tile);
        else hipLaunchKernelGGL((k_visc_walk<false, false>), g, kBlock, 0, st, P, tile);
    }
}

void launch_visc_fused(const KParams &P, hipStream_t st)
{
    const int tile = visc_tile(P);
    if (tile == 1) LAUNCH(k_visc_fused, cell_grid(P, P.nr + 1), kBlock, st, P);
    else launch_visc_walk(P, tile, false, st);
}

void launch_visc_energy(const KParams &P, hipStream_t st)
{
    const int tile = visc_tile(P);
    if (tile == 1) {
        LAUNCH(k_visc_fused, cell_grid(P, P.nr + 1), kBlock, st, P);
        launch_substep3(P, st);
    } else {
        launch_visc_walk(P, tile, true, st);
    }
}

// rings per thread of the kick walk: 8, or 4, while that gives 3072
// wavefronts, else 1 (the separate passes)
static int kick_tile(const KParams &P)
{
    if (P.kick_tile > 0) return P.kick_tile;
    const long segs = (P.ns + kKickSegW - 1) / kKickSegW;
    int tile = 8;
    while (tile > 1 && segs * ((P.nr + tile - 1) / tile) < 3072) tile = tile > 4 ? tile / 2 : 1;
    return tile;
}

int resolve_kick_tile(const KParams &P) { return kick_tile(P); }
int resolve_visc_tile(const KParams &P) { return visc_tile(P); }
long walk_waves(const KParams &P, int tile)
{
    const long rings = (P.nr + tile - 1) / tile;
    return rings * ((P.ns + kKickSegW - 1) / kKickSegW + (P.ns + kViscSegW - 1) / kViscSegW) / 2;
}

bool kick_walk_applies(const KParams &P)
{
    return (P.artvisc == FCPT_ARTVISC_TW || P.artvisc == FCPT_ARTVISC_SN) && kick_tile(P) > 1 &&
           visc_tile(P) > 1;
}

void launch_kick(const KParams &P, hipStream_t st, int fold_mode)
{
    const int tile = kick_tile(P);
    const int tiles = (P.nr + tile - 1) / tile;
    const dim3 g((unsigned)((P.ns + kKickSegW - 1) / kKickSegW), (unsigned)((tiles + 3) / 4));
    launch_potential_fold(P, st, fold_mode);
    const int v = (P.split ? 4 : 0) + (P.adiabatic ? 2 : 0) + (P.artvisc == FCPT_ARTVISC_SN ? 1 : 0);
    switch (v) {
    case 0: hipLaunchKernelGGL((k_kick_walk<false, false, false>), g, kBlock, 0, st, P, tile); break;
    case 1: hipLaunchKernelGGL((k_kick_walk<false, false, true>), g, kBlock, 0, st, P, tile); break;
    case 2: hipLaunchKernelGGL((k_kick_walk<false, true, false>), g, kBlock, 0, st, P, tile); break;
    case 3: hipLaunchKernelGGL((k_kick_walk<false, true, true>), g, kBlock, 0, st, P, tile); break;
    case 4: hipLaunchKernelGGL((k_kick_walk<true, false, false>), g, kBlock, 0, st, P, tile); break;
    case 5: hipLaunchKernelGGL((k_kick_walk<true, false, true>), g, kBlock, 0, st, P, tile); break;
    case 6: hipLaunchKernelGGL((k_kick_walk<true, true, false>), g, kBlock, 0, st, P, tile); break;
    default: hipLaunchKernelGGL((k_kick_walk<true, true, true>), g, kBlock, 0, st, P, tile); break;
    }
}

void launch_derived(const KParams &P, hipStream_t st) { LAUNCH(k_derived, cell_grid(P, P.nr), kBlock, st, P); }

// the final boundary call (damping, ghost rings, the clock) and the derived
// fields: damping, then one pass for the rest
void launch_final_boundaries_derived(const KParams &P, hipStream_t st, int advance_time)
{
    if (P.damping) {
        const int n_inner = (P.damp_ids[0] > P.damp_ids[1] ? P.damp_ids[0] : P.damp_ids[1]) + 1;
        const int o0 = P.damp_ids[2] < P.damp_ids[3] ? P.damp_ids[2] : P.damp_ids[3];
        const int n_outer = P.nr + 1 - o0;
        LAUNCH(k_damping, cell_grid(P, n_inner + n_outer), kBlock, st, P);
    }
    LAUNCH(k_final_derived, cell_grid(P, P.nr), kBlock, st, P, advance_time);
}
void launch_iso_rings(const KParams &P, hipStream_t st)
{
    hipLaunchKernelGGL(k_iso_rings, dim3((unsigned)((P.nr + 255) / 256)), dim3(256), 0, st, P);
}

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

void launch_split_ring_mean(const KParams &P, hipStream_t st, const void *field, const double *bg,
                            int rows, double *out)
{
    hipLaunchKernelGGL(k_split_ring_mean, dim3((unsigned)rows), dim3(256), 0, st,
                       static_cast<const float *>(field), bg, P.ns, out);
}

void launch_totals(const KParams &P, hipStream_t st, double radius, double *ring_scratch,
                   double *out2)
{
    LAUNCH(k_totals_rings, dim3(P.nr), dim3(64), st, P, radius, ring_scratch);
    hipLaunchKernelGGL(k_totals_final, dim3(1), dim3(1), 0, st, P, ring_scratch, out2);
}

// ---------------------------------------------------------------------------
// The disk's gravity on the N bodies (ComputeDiskOnPlanetAccel, Force.cpp:23-122)

// ComputeAverageDensity, Pframeforce.cpp:174-187: ring means of Sigma over the
// active rings (pairwise tree order), one block per ring
template <bool SPLIT>
__global__ void __launch_bounds__(256) k_sigma_ring_mean(KParams P, double *sig1d)
{
    __shared__ double sh[256];
    const int ns = P.ns;
    const int i = blockIdx.x;
    FLD(S, FCPT_SIGMA);
    double acc = 0.0;
    strided_row<256>(ns, (int)threadIdx.x, [&](int j) { return S(i, j); }, [&](double x) { acc += x; });
    const double sum = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) sig1d[i] = sum / ns;
}

// Per cell and body, the reference's expression (the smoothing of
// compute_smoothing, Force.cpp:145-159, chosen by the flags; the Klahr-Kley
// factor 4q^3 - 3q^4 inside r_cubic); inner (Rmed < r_body) and outer sums
// kept apart as in the reference.  NB bodies per launch (a template argument,
// so only their accumulators live in registers).  A thread owns one sector
// and walks the rings with a grid stride, loading the next ring's cell before
// it works on the current one; 1 / d^3 comes from the reciprocal square root
// with two Newton steps (a few ulp from the reference's sqrt and division).
// Each block folds its 4 NB partial sums with wavefront shuffles and LDS into
// its own slot (component-major, slot = block), and k_disk_force_final sums
// each component's slots in a fixed order: deterministic, no atomics.
template <bool SPLIT, int NB>
__global__ void __launch_bounds__(256) k_disk_force(KParams P, ForceBodies B, const double *sig1d,
                                                    double *part)
{
    const int ns = P.ns;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int nslots = (int)(gridDim.x * gridDim.y);
    double acc[NB][4];
#pragma unroll
    for (int k = 0; k < NB; ++k) acc[k][0] = acc[k][1] = acc[k][2] = acc[k][3] = 0.0;
    FLD(S, FCPT_SIGMA);
    FLD(Hf, FCPT_SCALE_HEIGHT);
    const double G = P.G;
    const double thick = P.thickness_smoothing;
    const bool correct = (B.flags & FCPT_FORCE_CORRECT_SELFGRAVITY) != 0;
    const int last = P.nr - 1;  // rings 1 .. nr - 2
    const int stride = P.cfl_gy * (int)blockDim.y;
    int i = wave_ring(1, (int)blockIdx.y);
    if (j < ns && i < last) {
        const double cphi = P.cosphi[j];
        const double sphi = P.sinphi[j];
        double sig = S(i, j);
        double H = P.adiabatic ? Hf(i, j) : 0.0;
        for (;;) {
            const int in = i + stride;
            const int ip = in < last ? in : i;
            const double sig_n = S(ip, j);
            const double H_n = P.adiabatic ? Hf(ip, j) : 0.0;
            const double rmed = cst(P.Rmed)[i];
            const double Hc = P.adiabatic ? H : cst(P.iso_h)[i];
            const double xc = rmed * cphi;
            const double yc = rmed * sphi;
            const double cell_sigma = correct ? sig - cst(sig1d)[i] : sig;
            const double gm = G * (cst(P.Surf)[i] * cell_sigma);
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const double smooth = B.smooth_fixed[k] >= 0.0 ? B.smooth_fixed[k] : thick * Hc;
                const double dx = xc - B.x[k];
                const double dy = yc - B.y[k];
                const double d2 = (dx * dx + dy * dy) + smooth * smooth;
                const double inv_d = inv_sqrt<true>(d2);
                const double inv_d3 = inv_d * inv_d * inv_d;
                double fx = gm * dx * inv_d3;
                double fy = gm * dy * inv_d3;
                const double r_sm = B.r_cubic[k];
                if (r_sm > 0.0) {  // uniform: the body's cubic smoothing radius
                    const double q = (d2 * inv_d) / r_sm;
                    const double klahr = q < 1.0 ? q * q * q * (4.0 - 3.0 * q) : 1.0;
                    fx *= klahr;
                    fy *= klahr;
                }
                if (rmed < B.r[k]) {  // uniform: one ring against one body
                    acc[k][0] += fx;
                    acc[k][1] += fy;
                } else {
                    acc[k][2] += fx;
                    acc[k][3] += fy;
                }
            }
            if (in >= last) break;
            i = in;
            sig = sig_n;
            H = H_n;
        }
    }
    __shared__ double wsum[4][NB * 4];
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double v = acc[k][c];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if ((t & 63) == 0) wsum[t >> 6][4 * k + c] = v;
        }
    }
    __syncthreads();
    if (t < NB * 4)
        part[(size_t)t * nslots + (blockIdx.y * gridDim.x + blockIdx.x)] =
            (wsum[0][t] + wsum[1][t]) + (wsum[2][t] + wsum[3][t]);
}

// one block per component: the slots summed with a 256-wide stride, then a
// pairwise tree over the 256 partial sums
__global__ void __launch_bounds__(256) k_disk_force_final(const double *part, int nslots, double *out)
{
    __shared__ double sh[256];
    const double *p = part + (size_t)blockIdx.x * nslots;
    double s = 0.0;
    for (int b = threadIdx.x; b < nslots; b += 256) s += p[b];
    s = block_sum<256>(s, sh);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

template <bool SPLIT, int NB>
static void launch_force_nb(const KParams &P, hipStream_t st, const ForceBodies &B, const double *sig1d,
                            double *part, dim3 g)
{
    hipLaunchKernelGGL((k_disk_force<SPLIT, NB>), g, kBlock, 0, st, P, B, sig1d, part);
}

template <bool SPLIT>
static void launch_force_split(const KParams &P, hipStream_t st, const ForceBodies &B,
                               const double *sig1d, double *part, dim3 g)
{
    switch (B.n) {
    case 1: launch_force_nb<SPLIT, 1>(P, st, B, sig1d, part, g); break;
    case 2: launch_force_nb<SPLIT, 2>(P, st, B, sig1d, part, g); break;
    case 3: launch_force_nb<SPLIT, 3>(P, st, B, sig1d, part, g); break;
    case 4: launch_force_nb<SPLIT, 4>(P, st, B, sig1d, part, g); break;
    case 5: launch_force_nb<SPLIT, 5>(P, st, B, sig1d, part, g); break;
    case 6: launch_force_nb<SPLIT, 6>(P, st, B, sig1d, part, g); break;
    case 7: launch_force_nb<SPLIT, 7>(P, st, B, sig1d, part, g); break;
    default: launch_force_nb<SPLIT, 8>(P, st, B, sig1d, part, g); break;
    }
}

void launch_disk_force(const KParams &P, hipStream_t st, const ForceBodies &B, double *sig1d,
                       double *part, double *out4)
{
    static_assert(kForceMaxBodies == 8, "launch_force_split dispatches 1 .. 8 bodies");
    if (B.n < 1 || B.n > kForceMaxBodies) return;
    if (B.flags & FCPT_FORCE_CORRECT_SELFGRAVITY)
        LAUNCH(k_sigma_ring_mean, dim3((unsigned)P.nr), dim3(256), st, P, sig1d);
    const dim3 g((unsigned)P.cfl_gx, (unsigned)P.cfl_gy);
    if (P.split) launch_force_split<true>(P, st, B, sig1d, part, g);
    else launch_force_split<false>(P, st, B, sig1d, part, g);
    hipLaunchKernelGGL(k_disk_force_final, dim3((unsigned)(4 * B.n)), dim3(256), 0, st, part,
                       (int)(g.x * g.y), out4);
}

// ---------------------------------------------------------------------------
// Batched step: many disks per launch.  Each stage of one sim::run iteration
// is one launch over a work table listing, for every disk that takes part in
// the stage, every block of that disk's own grid.  A disk's blocks do exactly
// what its single-disk launch does, so a batched disk evolves bit for bit like
// the same disk stepped alone.

static void add_grid(std::vector<BatchEntry> &v, unsigned d, unsigned gx, unsigned gy)
{
    for (unsigned y = 0; y < gy; ++y)
        for (unsigned x = 0; x < gx; ++x) v.push_back(BatchEntry{d, (y << 16) | x});
}

static unsigned rows_grid(const KParams &P, int rows) { return (unsigned)((rows + 3) / 4); }
static unsigned sector_grid(const KParams &P) { return (unsigned)((P.ns + 63) / 64); }

std::vector<BatchStageHost> plan_batch_step(const std::vector<KParams> &D)
{
    std::vector<BatchStageHost> st;
    // parameter block of disk d: d (the fields hold the velocities) or n + d
    // (its swapped view: the partner buffers hold them, between fused passes)
    const unsigned n = (unsigned)D.size();
    std::vector<unsigned> view(n, 0);
    auto stage_v = [&](int kind, int arg, auto pred, auto grid, bool flips, bool fields) {
        BatchStageHost s{kind, arg, {}};
        for (unsigned d = 0; d < n; ++d) {
            if (!pred(D[d])) continue;
            grid(s.entries, (fields || !view[d]) ? d : n + d, D[d]);
            if (flips) view[d] ^= 1u;
        }
        if (!s.entries.empty()) st.push_back(std::move(s));
    };
    auto stage = [&](int kind, int arg, auto pred, auto grid) { stage_v(kind, arg, pred, grid, false, false); };
    auto all = [](const KParams &) { return true; };
    auto adi = [](const KParams &P) { return P.adiabatic != 0; };
    auto iso = [](const KParams &P) { return P.adiabatic == 0; };
    auto rows = [](int off) {
        return [off](std::vector<BatchEntry> &v, unsigned d, const KParams &P) {
            add_grid(v, d, sector_grid(P), rows_grid(P, P.nr - off));
        };
    };
    auto rings = [](std::vector<BatchEntry> &v, unsigned d, const KParams &P) {
        add_grid(v, d, (unsigned)P.nr, 1);
    };
    auto cfl = [](std::vector<BatchEntry> &v, unsigned d, const KParams &P) {
        add_grid(v, d, (unsigned)P.cfl_gx, (unsigned)P.cfl_gy);
    };
    auto ghosts = [](std::vector<BatchEntry> &v, unsigned d, const KParams &P) {
        add_grid(v, d, (unsigned)((P.ns + 255) / 256), 1);
    };
    auto radial = [](std::vector<BatchEntry> &v, unsigned d, const KParams &P) {
        add_grid(v, d, sector_grid(P), (unsigned)((P.nr + P.radial_tile - 1) / P.radial_tile));
    };
    auto azim = [](std::vector<BatchEntry> &v, unsigned d, const KParams &P) {
        const int waves = (P.ns + kAzSegW - 1) / kAzSegW;
        add_grid(v, d, (unsigned)((waves + kAzWaves - 1) / kAzWaves), (unsigned)P.nr);
    };
    auto damp = [](std::vector<BatchEntry> &v, unsigned d, const KParams &P) {
        const int n_inner = (P.damp_ids[0] > P.damp_ids[1] ? P.damp_ids[0] : P.damp_ids[1]) + 1;
        const int o0 = P.damp_ids[2] < P.damp_ids[3] ? P.damp_ids[2] : P.damp_ids[3];
        add_grid(v, d, sector_grid(P), rows_grid(P, n_inner + P.nr + 1 - o0));
    };
    auto one = [](std::vector<BatchEntry> &v, unsigned d, const KParams &) { add_grid(v, d, 1, 1); };
    // sim::CalculateTimeStep and the monitor clip (simulation.cpp:100-118, 526-542):
    // the per-ring CFL blocks and the fold, as launch_cfl
    stage(B_CFL_RINGS, 0, all, rings);
    stage(B_CFL_FINAL_RINGS, 2, all, one);
    // step_Euler (simulation.cpp:148-267), in the order of enqueue_euler
    stage(B_POTENTIAL, 0, all, [](std::vector<BatchEntry> &v, unsigned d, const KParams &P) {
        add_grid(v, d, (unsigned)((P.ns + 64 * kPotCells - 1) / (64 * kPotCells)), rows_grid(P, P.nr));
    });
    // the kick: the two walks where the disk's own launch uses them
    // (kick_walk_applies with the tiles resolved at batch creation), else the
    // cell-parallel fused passes; each pass flips the disk's view
    auto flip = [&](int kind, int arg, auto pred, auto grid) { stage_v(kind, arg, pred, grid, true, false); };
    auto walk = [](const KParams &P) { return kick_walk_applies(P); };
    auto kick_grid = [](std::vector<BatchEntry> &v, unsigned d, const KParams &P) {
        const int tiles = (P.nr + P.kick_tile - 1) / P.kick_tile;
        add_grid(v, d, (unsigned)((P.ns + kKickSegW - 1) / kKickSegW), (unsigned)((tiles + 3) / 4));
    };
    auto visc_grid = [](std::vector<BatchEntry> &v, unsigned d, const KParams &P) {
        const int tiles = (P.nr + P.visc_tile - 1) / P.visc_tile;
        add_grid(v, d, (unsigned)((P.ns + kViscSegW - 1) / kViscSegW), (unsigned)((tiles + 3) / 4));
    };
    for (int a = 0; a < 2; ++a)
        for (int sn = 0; sn < 2; ++sn)
            flip(B_KICK_WALK, 2 * a + sn,
                 [&](const KParams &P) {
                     return walk(P) && P.adiabatic == a && (P.artvisc == FCPT_ARTVISC_SN) == (sn != 0);
                 },
                 kick_grid);
    auto cell = [&](const KParams &P) { return !walk(P); };
    flip(B_SRC_FUSED, 0, cell, rows(-1));
    stage(B_SRC_ENERGY, 0, [&](const KParams &P) { return cell(P) && P.adiabatic; }, rows(1));
    auto av = [&](const KParams &P) {
        return cell(P) && (P.artvisc == FCPT_ARTVISC_TW || P.artvisc == FCPT_ARTVISC_SN);
    };
    auto none = [](const KParams &P) { return P.artvisc == FCPT_ARTVISC_NONE; };
    flip(B_AV_FUSED, 1, av, rows(-1));
    stage(B_TEMP_FLOOR, 0,
          [](const KParams &P) {
              return P.artvisc == FCPT_ARTVISC_NONE && P.adiabatic && P.artvisc_dissipation;
          },
          rows(0));
    stage(B_RECALC_VISC, 0, none, rows(0));
    // the viscous pass: the walk (with SubStep3 when adiabatic) where the
    // disk's tile is above 1, as launch_visc_fused / launch_visc_energy
    auto vwalk = [](const KParams &P) { return P.visc_tile > 1; };
    for (int a = 0; a < 2; ++a)
        flip(B_VISC_WALK, a, [&](const KParams &P) { return vwalk(P) && P.adiabatic == a; }, visc_grid);
    flip(B_VISC_FUSED, 0, [&](const KParams &P) { return !vwalk(P); }, rows(-1));
    stage(B_SUBSTEP3, 0, [&](const KParams &P) { return !vwalk(P) && P.adiabatic; }, rows(0));
    // the boundary call before the transport with the FARGO ring means
    stage(B_GHOST_MEAN, 0, all, [](std::vector<BatchEntry> &v, unsigned d, const KParams &P) {
        add_grid(v, d, (unsigned)((P.ns + 255) / 256 + P.nr), 1);
    });
    stage(B_RADIAL_ISO, 0, iso, radial);
    stage(B_RADIAL_ADI, 0, adi, radial);
    stage(B_AZIM_ISO, 0, iso, azim);
    stage(B_AZIM_ADI, 0, adi, azim);
    // the velocity kernel with the final boundary call and the derived
    // fields writes the fields of every disk
    stage_v(B_VEL_FINAL, 2, all, rows(0), false, true);
    std::fill(view.begin(), view.end(), 0u);
    (void)ghosts;
    (void)cfl;
    (void)damp;
    return st;
}

#define LAUNCHB(kern, block, ...)                                                              \
    do {                                                                                       \
        if (split) hipLaunchKernelGGL(kern<true>, dim3(S.n), block, 0, st, Ps, S.map, ##__VA_ARGS__);  \
        else hipLaunchKernelGGL(kern<false>, dim3(S.n), block, 0, st, Ps, S.map, ##__VA_ARGS__);       \
    } while (0)

void launch_batch_stage(const KParams *Ps, const BatchStageDev &S, int split, int radial_waves,
                        hipStream_t st)
{
    const dim3 b256(256);
    switch (S.kind) {
    case B_RING_MEAN: LAUNCHB(kb_ring_mean, b256, S.arg); break;
    case B_CFL_CELLS: LAUNCHB(kb_cfl_cells, kBlock); break;
    case B_CFL_FINAL:
        hipLaunchKernelGGL(kb_cfl_finalize, dim3(S.n), b256, 0, st, Ps, S.map, S.arg);
        break;
    case B_POTENTIAL: LAUNCHB(kb_potential, kBlock); break;
    case B_SRC_VR: LAUNCHB(kb_src_vr, kBlock); break;
    case B_SRC_VPHI: LAUNCHB(kb_src_vphi, kBlock); break;
    case B_SRC_FUSED: LAUNCHB(kb_src_fused, kBlock); break;
    case B_AV_FUSED: LAUNCHB(kb_av_fused, kBlock, S.arg); break;
    case B_VISC_FUSED: LAUNCHB(kb_visc_fused, kBlock); break;
    case B_SRC_ENERGY: LAUNCHB(kb_src_energy, kBlock); break;
    case B_TW_Q: LAUNCHB(kb_tw_q, kBlock); break;
    case B_TW_VEL: LAUNCHB(kb_tw_vel, kBlock, S.arg); break;
    case B_SN_Q: LAUNCHB(kb_sn_q, kBlock); break;
    case B_SN_VEL: LAUNCHB(kb_sn_vel, kBlock, S.arg); break;
    case B_TEMP_FLOOR: LAUNCHB(kb_temperature_floor, kBlock); break;
    case B_RECALC_VISC: LAUNCHB(kb_recalc_viscosity, kBlock); break;
    case B_STRESS: LAUNCHB(kb_stress, kBlock); break;
    case B_VISC_UPDATE: LAUNCHB(kb_visc_update, kBlock); break;
    case B_SUBSTEP3: LAUNCHB(kb_substep3, kBlock); break;
    case B_GHOST: LAUNCHB(kb_ghost_bc, b256, S.arg); break;
    case B_VELOCITIES: LAUNCHB(kb_velocities, kBlock); break;
    case B_DAMPING: LAUNCHB(kb_damping, kBlock); break;
    case B_DERIVED: LAUNCHB(kb_derived, kBlock); break;
    case B_FINAL_DERIVED: LAUNCHB(kb_final_derived, kBlock, S.arg); break;
    case B_CFL_RINGS: LAUNCHB(kb_cfl_rings, dim3(512)); break;
    case B_CFL_FINAL_RINGS:
        hipLaunchKernelGGL(kb_cfl_finalize_rings, dim3(S.n), b256, 0, st, Ps, S.map, S.arg);
        break;
    case B_GHOST_MEAN: LAUNCHB(kb_ghost_ring_mean, b256); break;
    case B_VEL_FINAL: LAUNCHB(kb_velocities_final, kBlock, S.arg); break;
    case B_KICK_WALK: {
        const dim3 g(S.n);
#define KWB(SP, AD, SN)                                                                            hipLaunchKernelGGL((kb_kick_walk<SP, AD, SN>), g, kBlock, 0, st, Ps, S.map)
        const int v = (split ? 4 : 0) + S.arg;
        if (v == 0) KWB(false, false, false);
        else if (v == 1) KWB(false, false, true);
        else if (v == 2) KWB(false, true, false);
        else if (v == 3) KWB(false, true, true);
        else if (v == 4) KWB(true, false, false);
        else if (v == 5) KWB(true, false, true);
        else if (v == 6) KWB(true, true, false);
        else KWB(true, true, true);
#undef KWB
        break;
    }
    case B_VISC_WALK: {
        const dim3 g(S.n);
        if (split) {
            if (S.arg) hipLaunchKernelGGL((kb_visc_walk<true, true>), g, kBlock, 0, st, Ps, S.map);
            else hipLaunchKernelGGL((kb_visc_walk<true, false>), g, kBlock, 0, st, Ps, S.map);
        } else {
            if (S.arg) hipLaunchKernelGGL((kb_visc_walk<false, true>), g, kBlock, 0, st, Ps, S.map);
            else hipLaunchKernelGGL((kb_visc_walk<false, false>), g, kBlock, 0, st, Ps, S.map);
        }
        break;
    }
    case B_RADIAL_ISO:
    case B_RADIAL_ADI: {
        const bool a = S.kind == B_RADIAL_ADI, w2 = radial_waves == 2;
        const dim3 g(S.n), b(64);
#define RADB(SP, AD)                                                                           \
    do {                                                                                       \
        if (w2) hipLaunchKernelGGL((kb_radial_sweep2<SP, AD>), g, b, 0, st, Ps, S.map);        \
        else hipLaunchKernelGGL((kb_radial_sweep3<SP, AD>), g, b, 0, st, Ps, S.map);           \
    } while (0)
        if (split) {
            if (a) RADB(true, true);
            else RADB(true, false);
        } else {
            if (a) RADB(false, true);
            else RADB(false, false);
        }
#undef RADB
        break;
    }
    case B_AZIM_ISO:
    case B_AZIM_ADI: {
        const bool a = S.kind == B_AZIM_ADI;
        const dim3 g(S.n), b(64 * kAzWaves);
        if (split) {
            if (a) hipLaunchKernelGGL((kb_azimuthal_sweep<true, true>), g, b, 0, st, Ps, S.map);
            else hipLaunchKernelGGL((kb_azimuthal_sweep<true, false>), g, b, 0, st, Ps, S.map);
        } else {
            if (a) hipLaunchKernelGGL((kb_azimuthal_sweep<false, true>), g, b, 0, st, Ps, S.map);
            else hipLaunchKernelGGL((kb_azimuthal_sweep<false, false>), g, b, 0, st, Ps, S.map);
        }
        break;
    }
    default: break;
    }
}

}  // namespace fcpt
