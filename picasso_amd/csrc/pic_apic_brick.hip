// APIC brick scheme for particles binned by pic_bin_sort.
//
// A workgroup processes bricks of BB^3 bins (looping over the list of
// non-empty bricks when the sort provided one).  The bins are aligned with
// the spline's knot base (node-centred dual cells for quadratic splines on
// nodes, mesh cells for linear/cubic), so every particle of a bin shares one
// stencil and a brick's particles touch a fixed (BB+order)^3 node footprint.
//
// p2g: the footprint is accumulated in LDS.  Each work-item owns bins (not
//   particles): it sums its bin's particle contributions in registers (FP32
//   orders 1-2) or adds them straight into LDS, then adds the bin sums into
//   the LDS footprint.  FP32 LDS sums are compensated: the returned old value
//   of the LDS atomic gives the exact rounding error of each add (two-sum),
//   which is accumulated in a second LDS array and folded in at the flush.
//   The footprint is flushed to HBM with one global atomic per touched node
//   and component (nodes shared with neighbouring bricks receive one partial
//   per brick).
// g2p: the footprint of grid values is staged in LDS with row loads (from
//   the velocity, or from momentum / mass when normalisation is fused in);
//   each work-item then reconstructs one particle from LDS.
// A particle whose stencil is not where its bin says (positions changed
// since the sort) takes a slow path straight to global memory, so results
// never depend on the bins being current.
#include "pic_apic.hpp"
#include "pic_sort.hpp"

namespace pic {

constexpr int BB = 8;               // bins per brick edge
constexpr int BVOL = BB * BB * BB;  // bins per brick
constexpr int BT = 256;             // work-items per brick workgroup
constexpr int MAX_BRICK_WG = 2048;  // workgroups looping over the active list

template <typename T>
struct BrickArgs {
    ApicArgs<T> a;
    const uint32_t* keys;
    const uint32_t* bin_start;
    const uint32_t* active;  // [count, ids...] of non-empty bricks, or null
    uint32_t nbricks;
    int32_t cell_lo[3];
    int32_t nb[3];
    int32_t knot_off;  // first knot index = bin index + knot_off
};

template <typename T>
__device__ __forceinline__ void brick_origin(const BrickArgs<T>& g, uint32_t b, int f0[3])
{
    const int bk = int(b % uint32_t(g.nb[2]));
    const int bj = int((b / uint32_t(g.nb[2])) % uint32_t(g.nb[1]));
    const int bi = int(b / (uint32_t(g.nb[2]) * uint32_t(g.nb[1])));
    f0[0] = g.cell_lo[0] + bi * BB + g.knot_off;
    f0[1] = g.cell_lo[1] + bj * BB + g.knot_off;
    f0[2] = g.cell_lo[2] + bk * BB + g.knot_off;
}

template <typename T>
__device__ __forceinline__ uint32_t brick_count(const BrickArgs<T>& g)
{
    return g.active ? g.active[0] : g.nbricks;
}

template <typename T>
__device__ __forceinline__ uint32_t brick_id(const BrickArgs<T>& g, uint32_t w)
{
    return g.active ? g.active[1 + w] : w;
}

// LDS accumulation; FP32 keeps the exact rounding error of every add.
__device__ __forceinline__ void lds_acc(float* v, float* c, float x)
{
    const float old = atomicAdd(v, x);
    const float s = old + x;
    const float bb = s - old;
    const float e = (old - (s - bb)) + (x - bb);
    atomicAdd(c, e);
}
__device__ __forceinline__ void lds_acc(double* v, double*, double x) { atomicAdd(v, x); }

// Slow path: one particle straight to the global grid.
template <int ORDER, typename T>
__device__ void p2g_global(const ApicArgs<T>& a, const ApicParticle<T>& pt,
                           const Spline<ORDER, T>& sd)
{
    if (!stencil_inside(sd, a.gmu)) {
        atomicOr(a.err, ERRBIT_BOUNDS);
        return;
    }
    apic_p2g_particle<ORDER, false>(a, pt, sd, [&](int i, int j, int k, const T* mu, T m) {
        const int64_t off = a.gmu.off(sd.s0[0] + i, sd.s0[1] + j, sd.s0[2] + k);
#pragma unroll
        for (int d = 0; d < 3; ++d) atomic_add(a.gmu.data + off + d * a.gmu.cs, mu[d]);
        atomic_add(a.gm.data + a.gm.off(sd.s0[0] + i, sd.s0[1] + j, sd.s0[2] + k), m);
    });
}

template <int ORDER, typename T, bool REGACC>
__device__ __forceinline__ void p2g_brick(const BrickArgs<T>& g, uint32_t b, T* accv, T* accc,
                                          uint32_t* cst)
{
    constexpr int NK = ORDER + 1;
    constexpr int NN = NK * NK * NK;
    constexpr int FN = BB + ORDER;
    constexpr int FV = FN * FN * FN;
    constexpr bool COMP = std::is_same<T, float>::value;

    const uint32_t beg = g.bin_start[b], end = g.bin_start[b + 1];
    if (beg == end) return;
    int f0[3];
    brick_origin(g, b, f0);

    for (int q = threadIdx.x; q < 4 * FV; q += BT) {
        accv[q] = T(0);
        if constexpr (COMP) accc[q] = T(0);
    }
    for (int q = threadIdx.x; q <= BVOL; q += BT) cst[q] = 0u;
    __syncthreads();
    // Bin ranges inside the brick from the sorted keys.
    const uint32_t kb = b * BVOL;
    for (uint32_t s = beg + threadIdx.x; s < end; s += BT) atomicAdd(&cst[g.keys[s] - kb + 1], 1u);
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        uint32_t carry = 0;
        for (int c0 = 1; c0 <= BVOL; c0 += 64) {
            const uint32_t v = (c0 + lane <= BVOL) ? cst[c0 + lane] : 0u;
            uint32_t inc = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(inc, o);
                if (lane >= o) inc += y;
            }
            if (c0 + lane <= BVOL) cst[c0 + lane] = carry + inc;
            carry += __shfl(inc, 63);
        }
    }
    __syncthreads();

    for (int l = threadIdx.x; l < BVOL; l += BT) {
        const uint32_t s0 = beg + cst[l], s1 = beg + cst[l + 1];
        if (s0 == s1) continue;
        const int li = l / (BB * BB), lj = (l / BB) % BB, lk = l % BB;
        const int e0 = f0[0] + li, e1 = f0[1] + lj, e2 = f0[2] + lk;
        const int lbase = (li * FN + lj) * FN + lk;
        T acc[REGACC ? NN : 1][4];
        if constexpr (REGACC) {
#pragma unroll
            for (int n = 0; n < NN; ++n) acc[n][0] = acc[n][1] = acc[n][2] = acc[n][3] = T(0);
        }
        for (uint32_t s = s0; s < s1; ++s) {
            const int64_t p = int64_t(g.a.perm[s]);
            ApicParticle<T> pt;
            load_apic_particle<false>(g.a, p, pt);
            Spline<ORDER, T> sd;
            eval_spline<ORDER, T, ORDER == 1, ORDER != 1>(g.a.lat, pt.x, sd);
            if (sd.s0[0] != e0 || sd.s0[1] != e1 || sd.s0[2] != e2) {
                p2g_global<ORDER, T>(g.a, pt, sd);
                continue;
            }
            apic_p2g_particle<ORDER, false>(g.a, pt, sd, [&](int i, int j, int k, const T* mu, T m) {
                if constexpr (REGACC) {
                    T* a4 = acc[(i * NK + j) * NK + k];
                    a4[0] += m;
                    a4[1] += mu[0];
                    a4[2] += mu[1];
                    a4[3] += mu[2];
                } else {
                    const int q = lbase + (i * FN + j) * FN + k;
                    lds_acc(&accv[q], &accc[COMP ? q : 0], m);
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        lds_acc(&accv[(d + 1) * FV + q], &accc[COMP ? (d + 1) * FV + q : 0], mu[d]);
                }
            });
        }
        if constexpr (REGACC) {
#pragma unroll
            for (int i = 0; i < NK; ++i)
#pragma unroll
                for (int j = 0; j < NK; ++j)
#pragma unroll
                    for (int k = 0; k < NK; ++k) {
                        const int n = (i * NK + j) * NK + k;
                        const int q = lbase + (i * FN + j) * FN + k;
#pragma unroll
                        for (int c = 0; c < 4; ++c)
                            lds_acc(&accv[c * FV + q], &accc[COMP ? c * FV + q : 0], acc[n][c]);
                    }
        }
    }
    __syncthreads();

    // Flush the footprint.
    for (int q = threadIdx.x; q < FV; q += BT) {
        const int fi = q / (FN * FN), fj = (q / FN) % FN, fk = q % FN;
        T v[4];
        bool any = false;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            v[c] = accv[c * FV + q];
            if constexpr (COMP) v[c] += accc[c * FV + q];
            any = any || v[c] != T(0);
        }
        if (!any) continue;
        const int ni = f0[0] + fi, nj = f0[1] + fj, nk = f0[2] + fk;
        if (!(g.a.gmu.covers(0, ni, ni) && g.a.gmu.covers(1, nj, nj) && g.a.gmu.covers(2, nk, nk))) {
            atomicOr(g.a.err, ERRBIT_BOUNDS);
            continue;
        }
        atomic_add(g.a.gm.data + g.a.gm.off(ni, nj, nk), v[0]);
        const int64_t off = g.a.gmu.off(ni, nj, nk);
#pragma unroll
        for (int d = 0; d < 3; ++d) atomic_add(g.a.gmu.data + off + d * g.a.gmu.cs, v[d + 1]);
    }
}

template <int ORDER, typename T, bool REGACC>
__global__ __launch_bounds__(BT) void k_apic_p2g_brick(BrickArgs<T> g)
{
    constexpr int FN = BB + ORDER;
    constexpr int FV = FN * FN * FN;
    constexpr bool COMP = std::is_same<T, float>::value;
    __shared__ T accv[4 * FV];
    __shared__ T accc[COMP ? 4 * FV : 1];
    __shared__ uint32_t cst[BVOL + 1];
    const uint32_t nw = brick_count(g);
    for (uint32_t w = blockIdx.x; w < nw; w += gridDim.x) {
        p2g_brick<ORDER, T, REGACC>(g, brick_id(g, w), accv, accc, cst);
        __syncthreads();
    }
}

template <int ORDER, typename T>
__device__ __forceinline__ void g2p_brick(const BrickArgs<T>& g, uint32_t b, T* us)
{
    constexpr int FN = BB + ORDER;
    constexpr int FV = FN * FN * FN;
    const uint32_t beg = g.bin_start[b], end = g.bin_start[b + 1];
    if (beg == end) return;
    int f0[3];
    brick_origin(g, b, f0);
    const GView<T>& gu = g.a.gmu;
    for (int q = threadIdx.x; q < FV; q += BT) {
        const int ni = f0[0] + q / (FN * FN), nj = f0[1] + (q / FN) % FN, nk = f0[2] + q % FN;
        const bool in = gu.covers(0, ni, ni) && gu.covers(1, nj, nj) && gu.covers(2, nk, nk);
        const int64_t off = in ? gu.off(ni, nj, nk) : 0;
        const int64_t moff = in && g.a.gnm.data ? g.a.gnm.off(ni, nj, nk) : 0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
            us[d * FV + q] = in ? node_velocity(gu, g.a.gnm, off, moff, d) : T(0);
    }
    __syncthreads();

    for (uint32_t s = beg + threadIdx.x; s < end; s += BT) {
        const int64_t p = int64_t(g.a.perm[s]);
        const int64_t xb = g.a.x.base(p);
        T xp[3] = {g.a.x.at(xb, 0), g.a.x.at(xb, 1), g.a.x.at(xb, 2)};
        Spline<ORDER, T> sd;
        eval_spline<ORDER, T, false, true>(g.a.lat, xp, sd);
        const int li = sd.s0[0] - f0[0], lj = sd.s0[1] - f0[1], lk = sd.s0[2] - f0[2];
        T cp[12];
        if (li >= 0 && li < BB && lj >= 0 && lj < BB && lk >= 0 && lk < BB) {
            const int lbase = (li * FN + lj) * FN + lk;
            apic_g2p_particle<ORDER, false>(g.a, sd, [&](int i, int j, int k, T* ue) {
                const int q = lbase + (i * FN + j) * FN + k;
                ue[0] = us[q];
                ue[1] = us[FV + q];
                ue[2] = us[2 * FV + q];
            }, cp);
        } else {
            if (!stencil_inside(sd, gu)) {
                atomicOr(g.a.err, ERRBIT_BOUNDS);
                continue;
            }
            apic_g2p_particle<ORDER, false>(g.a, sd, [&](int i, int j, int k, T* ue) {
                const int ni = sd.s0[0] + i, nj = sd.s0[1] + j, nk = sd.s0[2] + k;
                const int64_t off = gu.off(ni, nj, nk);
                const int64_t moff = g.a.gnm.data ? g.a.gnm.off(ni, nj, nk) : 0;
#pragma unroll
                for (int d = 0; d < 3; ++d) ue[d] = node_velocity(gu, g.a.gnm, off, moff, d);
            }, cp);
        }
        const int64_t cb = g.a.c.base(p);
#pragma unroll
        for (int r = 0; r < 12; ++r) g.a.c.at(cb, r) = cp[r];
    }
}

template <int ORDER, typename T>
__global__ __launch_bounds__(BT) void k_apic_g2p_brick(BrickArgs<T> g)
{
    constexpr int FN = BB + ORDER;
    constexpr int FV = FN * FN * FN;
    __shared__ T us[3 * FV];
    const uint32_t nw = brick_count(g);
    for (uint32_t w = blockIdx.x; w < nw; w += gridDim.x) {
        g2p_brick<ORDER, T>(g, brick_id(g, w), us);
        __syncthreads();
    }
}

// Knot offset of the bins for (order, entity), or false when the bin
// lattice is not the spline's knot-base lattice.  In half cells: the base is
// floor(t_node - centred + order2) and the bin floor(t_node + dual).
static bool aligned_knot_offset(int order, int entity, int dual, int* knot_off)
{
    if (entity != PIC_NODE && entity != PIC_CELL) return false;
    const int halves = (order == 2 ? 1 : 0) - (entity == PIC_CELL ? 1 : 0) - (dual ? 1 : 0);
    if (halves != 0 && halves != -2) return false;
    *knot_off = halves / 2 + (order == 1 ? 0 : -1);
    return true;
}

template <typename T>
static bool make_brick_args(pic_context* ctx, int order, int entity, const pic_mesh& mesh,
                            int64_t np, const pic_pfield& x, const pic_pfield* m,
                            const pic_pfield& c, const pic_grid* gm, const pic_grid& gmu,
                            const pic_grid* gmass, const pic_bins& bins, BrickArgs<T>& g,
                            unsigned& nwg)
{
    for (int d = 0; d < 3; ++d)
        if (bins.geom.brick[d] != BB) return false;
    if (!bins.keys || !bins.bin_start) return false;
    if (!aligned_knot_offset(order, entity, bins.geom.dual, &g.knot_off)) return false;
    const int64_t nb = pic_bin_count(&bins.geom);
    if (nb <= 0) return false;
    g.a = make_args<T>(ctx, entity, mesh, np, &x, m, &c, gm, gmu, bins.perm);
    if (gmass) g.a.gnm = make_gview<T>(*gmass);
    g.keys = bins.keys;
    g.bin_start = bins.bin_start;
    g.active = bins.active;
    g.nbricks = uint32_t(nb);
    for (int d = 0; d < 3; ++d) {
        g.cell_lo[d] = bins.geom.cell_lo[d];
        g.nb[d] = (bins.geom.ncell[d] + BB - 1) / BB;
    }
    nwg = bins.active ? unsigned(std::min<int64_t>(nb, MAX_BRICK_WG)) : unsigned(nb);
    return true;
}

int apic_p2g_brick(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                   int64_t np, const pic_pfield& x, const pic_pfield& m, const pic_pfield& c,
                   const pic_grid& gm, const pic_grid& gmu, const pic_bins& bins)
{
    unsigned nwg = 0;
    hipStream_t st = ctx->stream;
    if (dtype == PIC_F32) {
        BrickArgs<float> g;
        if (!make_brick_args<float>(ctx, order, entity, mesh, np, x, &m, c, &gm, gmu, nullptr,
                                    bins, g, nwg))
            return PIC_ERR_UNSUPPORTED;
        TimedLaunch tl(ctx, "apic_p2g");
        if (order == 1) k_apic_p2g_brick<1, float, true><<<nwg, BT, 0, st>>>(g);
        else if (order == 2) k_apic_p2g_brick<2, float, true><<<nwg, BT, 0, st>>>(g);
        else k_apic_p2g_brick<3, float, false><<<nwg, BT, 0, st>>>(g);
    } else {
        BrickArgs<double> g;
        if (!make_brick_args<double>(ctx, order, entity, mesh, np, x, &m, c, &gm, gmu, nullptr,
                                     bins, g, nwg))
            return PIC_ERR_UNSUPPORTED;
        TimedLaunch tl(ctx, "apic_p2g");
        if (order == 1) k_apic_p2g_brick<1, double, false><<<nwg, BT, 0, st>>>(g);
        else if (order == 2) k_apic_p2g_brick<2, double, false><<<nwg, BT, 0, st>>>(g);
        else k_apic_p2g_brick<3, double, false><<<nwg, BT, 0, st>>>(g);
    }
    return launch_status();
}

int apic_g2p_brick(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                   int64_t np, const pic_pfield& x, const pic_grid& gu, const pic_grid* gmass,
                   const pic_pfield& c, const pic_bins& bins)
{
    unsigned nwg = 0;
    hipStream_t st = ctx->stream;
    if (dtype == PIC_F32) {
        BrickArgs<float> g;
        if (!make_brick_args<float>(ctx, order, entity, mesh, np, x, nullptr, c, nullptr, gu,
                                    gmass, bins, g, nwg))
            return PIC_ERR_UNSUPPORTED;
        TimedLaunch tl(ctx, "apic_g2p");
        if (order == 1) k_apic_g2p_brick<1, float><<<nwg, BT, 0, st>>>(g);
        else if (order == 2) k_apic_g2p_brick<2, float><<<nwg, BT, 0, st>>>(g);
        else k_apic_g2p_brick<3, float><<<nwg, BT, 0, st>>>(g);
    } else {
        BrickArgs<double> g;
        if (!make_brick_args<double>(ctx, order, entity, mesh, np, x, nullptr, c, nullptr, gu,
                                     gmass, bins, g, nwg))
            return PIC_ERR_UNSUPPORTED;
        TimedLaunch tl(ctx, "apic_g2p");
        if (order == 1) k_apic_g2p_brick<1, double><<<nwg, BT, 0, st>>>(g);
        else if (order == 2) k_apic_g2p_brick<2, double><<<nwg, BT, 0, st>>>(g);
        else k_apic_g2p_brick<3, double><<<nwg, BT, 0, st>>>(g);
    }
    return launch_status();
}

}  // namespace pic
