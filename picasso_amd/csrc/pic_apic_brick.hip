// APIC brick scheme for particles binned by pic_bin_sort.
//
// One workgroup per non-empty brick of BB^3 bins (BB = 4 or 8).  The bins
// are aligned with the spline's knot base (node-centred dual cells for
// quadratic splines on nodes, mesh cells for linear/cubic), so every
// particle of a bin shares one stencil and a brick's particles touch a fixed
// (BB+order)^3 node footprint.
//
// p2g: the brick's particles are staged through LDS in chunks with
//   coalesced loads (all loads of a chunk in flight at once).  KL
//   consecutive work-items then share a bin: each takes every KL-th staged
//   particle of it, sums the contributions in registers, the group reduces
//   its sums across lanes, and one lane adds the bin sums into the LDS node
//   footprint.  FP32 LDS sums are compensated: the returned old value of the
//   LDS atomic gives the exact rounding error of each add (two-sum), which is
//   accumulated in a second LDS array and folded in at the flush.  The
//   footprint is flushed to HBM with one global atomic per touched node and
//   component (nodes shared with neighbouring bricks receive one partial per
//   brick).
// g2p: the footprint of grid values is staged in LDS with row loads (from
//   the velocity, or from momentum / mass when normalisation is fused in);
//   each work-item then reconstructs one particle from LDS.
// A particle whose stencil is not where its bin says (positions changed
// since the sort) is appended to a list that the scatter/gather kernels
// process afterwards, so results never depend on the bins being current.
#include "pic_apic.hpp"
#include "pic_sort.hpp"

namespace pic {

constexpr int BT = 256;    // work-items per brick workgroup
constexpr int CHUNK = 512;  // particles staged in LDS at a time
constexpr int NPF = 16;     // staged scalars per particle: x(3), m, c_p(12)

template <typename T>
struct BrickArgs {
    ApicArgs<T> a;
    const uint32_t* keys;
    const uint32_t* bin_start;
    const uint32_t* active;  // [count, ids...] of non-empty bricks, or null
    uint32_t* stale;         // [count, slots...] of particles for the slow path
    uint32_t nbricks;
    int32_t cell_lo[3];
    int32_t nb[3];
    int32_t knot_off;  // first knot index = bin index + knot_off
};

template <int BB, typename T>
__device__ __forceinline__ void brick_origin(const BrickArgs<T>& g, uint32_t b, int f0[3])
{
    const int bk = int(b % uint32_t(g.nb[2]));
    const int bj = int((b / uint32_t(g.nb[2])) % uint32_t(g.nb[1]));
    const int bi = int(b / (uint32_t(g.nb[2]) * uint32_t(g.nb[1])));
    f0[0] = g.cell_lo[0] + bi * BB + g.knot_off;
    f0[1] = g.cell_lo[1] + bj * BB + g.knot_off;
    f0[2] = g.cell_lo[2] + bk * BB + g.knot_off;
}

// Brick handled by this workgroup, or false when it has nothing to do.
template <typename T>
__device__ __forceinline__ bool my_brick(const BrickArgs<T>& g, uint32_t* b)
{
    if (g.active) {
        if (blockIdx.x >= g.active[0]) return false;
        *b = g.active[1 + blockIdx.x];
    } else {
        *b = blockIdx.x;
    }
    return g.bin_start[*b] != g.bin_start[*b + 1];
}

template <typename T>
__device__ __forceinline__ void push_stale(const BrickArgs<T>& g, uint32_t s)
{
    const uint32_t at = atomicAdd(g.stale, 1u);
    g.stale[1 + at] = s;
}

// Exact rounding error of the FP32 sum old + x (two-sum; contraction cannot
// apply: there are no products).
__device__ __forceinline__ float two_sum_err(float old, float x)
{
    const float s = old + x;
    const float bb = s - old;
    return (old - (s - bb)) + (x - bb);
}

// Compensated LDS accumulation: value += x, comp += rounding error of that add.
__device__ __forceinline__ void lds_two_sum(float* v, float* c, float x)
{
    const float old = atomicAdd(v, x);
    atomicAdd(c, two_sum_err(old, x));
}

// Exclusive scan of cst[1..n] in place (cst[0] = 0 stays), by one wave.
__device__ __forceinline__ void scan_counts(uint32_t* cst, int n)
{
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        uint32_t carry = 0;
        for (int c0 = 1; c0 <= n; c0 += 64) {
            const uint32_t v = (c0 + lane <= n) ? cst[c0 + lane] : 0u;
            uint32_t inc = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(inc, o);
                if (lane >= o) inc += y;
            }
            if (c0 + lane <= n) cst[c0 + lane] = carry + inc;
            carry += __shfl(inc, 63);
        }
    }
}

template <int BB, int ORDER, typename T, bool REGACC>
__global__ __launch_bounds__(BT) void k_apic_p2g_brick(BrickArgs<T> g)
{
    constexpr int BVOL = BB * BB * BB;
    constexpr int NK = ORDER + 1;
    constexpr int NN = NK * NK * NK;
    constexpr int FN = BB + ORDER;
    constexpr int FV = FN * FN * FN;
    constexpr bool COMP = std::is_same<T, float>::value;
    constexpr int KL = REGACC ? 4 : 1;
    // Small bricks give every wave a private footprint copy: bin sums are
    // then added with plain LDS read-modify-writes (LDS float atomics run at
    // a small fraction of the LDS rate), and the copies are merged at the
    // flush.
    constexpr bool PRIV = REGACC && BB == 4;
    constexpr int NCOPY = PRIV ? BT / 64 : 1;
    __shared__ T accv[NCOPY * 4 * FV];
    __shared__ T accc[COMP ? NCOPY * 4 * FV : 1];
    __shared__ T pbuf[NPF * CHUNK];
    __shared__ uint32_t cst[BVOL + 1];

    uint32_t b;
    if (!my_brick(g, &b)) return;
    const uint32_t beg = g.bin_start[b], end = g.bin_start[b + 1];
    int f0[3];
    brick_origin<BB>(g, b, f0);

    for (int q = threadIdx.x; q < NCOPY * 4 * FV; q += BT) {
        accv[q] = T(0);
        if constexpr (COMP) accc[q] = T(0);
    }
    for (int q = threadIdx.x; q <= BVOL; q += BT) cst[q] = 0u;
    __syncthreads();
    // Bin ranges inside the brick (relative to beg) from the sorted keys.
    const uint32_t kb = b * BVOL;
    for (uint32_t s = beg + threadIdx.x; s < end; s += BT) atomicAdd(&cst[g.keys[s] - kb + 1], 1u);
    __syncthreads();
    scan_counts(cst, BVOL);
    __syncthreads();

    const ApicArgs<T>& a = g.a;
    for (uint32_t c0 = beg; c0 < end; c0 += CHUNK) {
        const uint32_t cn = min(uint32_t(CHUNK), end - c0);
        // Stage the chunk: consecutive work-items load consecutive particles.
        for (uint32_t i = threadIdx.x; i < cn; i += BT) {
            ApicParticle<T> pt;
            load_apic_particle<false>(a, int64_t(a.perm[c0 + i]), pt);
            pbuf[0 * CHUNK + i] = pt.x[0];
            pbuf[1 * CHUNK + i] = pt.x[1];
            pbuf[2 * CHUNK + i] = pt.x[2];
            pbuf[3 * CHUNK + i] = pt.m;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                pbuf[(4 + d) * CHUNK + i] = pt.c0[d];
#pragma unroll
                for (int q = 0; q < 3; ++q) pbuf[(7 + d * 3 + q) * CHUNK + i] = pt.B[d][q];
            }
        }
        __syncthreads();
        // Bins overlapping the chunk, KL work-items per bin.
        const uint32_t lo = c0 - beg, hi = lo + cn;
        const int lb = int(g.keys[c0] - kb), le = int(g.keys[c0 + cn - 1] - kb);
        for (int l0 = lb; l0 <= le; l0 += BT / KL) {
            const int l = l0 + int(threadIdx.x) / KL;
            const int ql = int(threadIdx.x) % KL;
            const bool mine = l <= le;
            const uint32_t s0 = mine ? max(cst[l], lo) : 0u;
            const uint32_t s1 = mine ? min(cst[l + 1], hi) : 0u;
            const int li = l / (BB * BB), lj = (l / BB) % BB, lk = l % BB;
            const int e0 = f0[0] + li, e1 = f0[1] + lj, e2 = f0[2] + lk;
            const int lbase = (li * FN + lj) * FN + lk;
            T acc[REGACC ? NN : 1][4];
            if constexpr (REGACC) {
#pragma unroll
                for (int n = 0; n < NN; ++n) acc[n][0] = acc[n][1] = acc[n][2] = acc[n][3] = T(0);
            }
            for (uint32_t s = s0 + ql; s < s1; s += KL) {
                const uint32_t i = s - lo;
                ApicParticle<T> pt;
                pt.x[0] = pbuf[0 * CHUNK + i];
                pt.x[1] = pbuf[1 * CHUNK + i];
                pt.x[2] = pbuf[2 * CHUNK + i];
                pt.m = pbuf[3 * CHUNK + i];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    pt.c0[d] = pbuf[(4 + d) * CHUNK + i];
#pragma unroll
                    for (int q = 0; q < 3; ++q) pt.B[d][q] = pbuf[(7 + d * 3 + q) * CHUNK + i];
                }
                Spline<ORDER, T> sd;
                eval_spline<ORDER, T, ORDER == 1, ORDER != 1>(a.lat, pt.x, sd);
                if (sd.s0[0] != e0 || sd.s0[1] != e1 || sd.s0[2] != e2) {
                    push_stale(g, beg + s);
                    continue;
                }
                apic_p2g_particle<ORDER, false>(a, pt, sd, [&](int i, int j, int k, const T* mu, T m) {
                    if constexpr (REGACC) {
                        T* a4 = acc[(i * NK + j) * NK + k];
                        a4[0] += m;
                        a4[1] += mu[0];
                        a4[2] += mu[1];
                        a4[3] += mu[2];
                    } else {
                        const int q = lbase + (i * FN + j) * FN + k;
                        if constexpr (COMP) {
                            lds_two_sum(&accv[q], &accc[q], m);
#pragma unroll
                            for (int d = 0; d < 3; ++d)
                                lds_two_sum(&accv[(d + 1) * FV + q], &accc[(d + 1) * FV + q], mu[d]);
                        } else {
                            atomicAdd(&accv[q], m);
#pragma unroll
                            for (int d = 0; d < 3; ++d) atomicAdd(&accv[(d + 1) * FV + q], mu[d]);
                        }
                    }
                });
            }
            if constexpr (REGACC) {
#pragma unroll
                for (int n = 0; n < NN; ++n)
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        static_assert(KL == 4, "bin groups are DPP quads");
                        acc[n][c] = quad_sum(acc[n][c]);
                    }
                if constexpr (PRIV) {
                    // Work-item ql of the group adds component ql into its
                    // wave's copy, one stencil node per step.  Within a step
                    // the wave's bins hit distinct nodes, and the steps stay
                    // in program order (the compiler may not move LDS
                    // accesses across the empty asm), so plain
                    // read-modify-writes are race-free and the two-sum sees
                    // the exact running value.
                    const int wv = int(threadIdx.x) / 64;
                    T* vp = accv + (wv * 4 + ql) * FV + lbase;
                    T* cp = accc + (COMP ? (wv * 4 + ql) * FV + lbase : 0);
                    if (s1 > s0) {
#pragma unroll
                        for (int n = 0; n < NN; ++n) {
                            const int i = n / (NK * NK), j = (n / NK) % NK, k = n % NK;
                            const int q = (i * FN + j) * FN + k;
                            const T x = ql == 0 ? acc[n][0]
                                                : (ql == 1 ? acc[n][1] : (ql == 2 ? acc[n][2] : acc[n][3]));
                            asm volatile("" ::: "memory");
                            const T old = vp[q];
                            vp[q] = old + x;
                            if constexpr (COMP) cp[q] += two_sum_err(old, x);
                        }
                    }
                    asm volatile("" ::: "memory");
                } else if (ql == 0 && s1 > s0) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        if constexpr (COMP) {
                            // All returning adds of a component are issued
                            // before any result is used, so their latencies
                            // overlap.
                            T old[NN];
#pragma unroll
                            for (int n = 0; n < NN; ++n) {
                                const int i = n / (NK * NK), j = (n / NK) % NK, k = n % NK;
                                old[n] = atomicAdd(&accv[c * FV + lbase + (i * FN + j) * FN + k],
                                                   acc[n][c]);
                            }
#pragma unroll
                            for (int n = 0; n < NN; ++n) {
                                const int i = n / (NK * NK), j = (n / NK) % NK, k = n % NK;
                                atomicAdd(&accc[c * FV + lbase + (i * FN + j) * FN + k],
                                          two_sum_err(old[n], acc[n][c]));
                            }
                        } else {
#pragma unroll
                            for (int n = 0; n < NN; ++n) {
                                const int i = n / (NK * NK), j = (n / NK) % NK, k = n % NK;
                                atomicAdd(&accv[c * FV + lbase + (i * FN + j) * FN + k], acc[n][c]);
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
    }

    // Flush the footprint (merging the per-wave copies, compensated).
    for (int q = threadIdx.x; q < FV; q += BT) {
        const int fi = q / (FN * FN), fj = (q / FN) % FN, fk = q % FN;
        T v[4];
        bool any = false;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            T s = accv[c * FV + q];
            T e = COMP ? accc[c * FV + q] : T(0);
#pragma unroll
            for (int w = 1; w < NCOPY; ++w) {
                const T x = accv[(w * 4 + c) * FV + q];
                if constexpr (COMP) e += two_sum_err(s, x) + accc[(w * 4 + c) * FV + q];
                s += x;
            }
            v[c] = s + e;
            any = any || v[c] != T(0);
        }
        if (!any) continue;
        const int ni = f0[0] + fi, nj = f0[1] + fj, nk = f0[2] + fk;
        if (!(a.gmu.covers(0, ni, ni) && a.gmu.covers(1, nj, nj) && a.gmu.covers(2, nk, nk))) {
            atomicOr(a.err, ERRBIT_BOUNDS);
            continue;
        }
        atomic_add(a.gm.data + a.gm.off(ni, nj, nk), v[0]);
        const int64_t off = a.gmu.off(ni, nj, nk);
#pragma unroll
        for (int d = 0; d < 3; ++d) atomic_add(a.gmu.data + off + d * a.gmu.cs, v[d + 1]);
    }
}

template <int BB, int ORDER, typename T>
__global__ __launch_bounds__(BT) void k_apic_g2p_brick(BrickArgs<T> g)
{
    constexpr int FN = BB + ORDER;
    constexpr int FV = FN * FN * FN;
    __shared__ T us[3 * FV];

    uint32_t b;
    if (!my_brick(g, &b)) return;
    const uint32_t beg = g.bin_start[b], end = g.bin_start[b + 1];
    int f0[3];
    brick_origin<BB>(g, b, f0);
    const ApicArgs<T>& a = g.a;
    const GView<T>& gu = a.gmu;
    for (int q = threadIdx.x; q < FV; q += BT) {
        const int ni = f0[0] + q / (FN * FN), nj = f0[1] + (q / FN) % FN, nk = f0[2] + q % FN;
        const bool in = gu.covers(0, ni, ni) && gu.covers(1, nj, nj) && gu.covers(2, nk, nk);
        const int64_t off = in ? gu.off(ni, nj, nk) : 0;
        const int64_t moff = in && a.gnm.data ? a.gnm.off(ni, nj, nk) : 0;
#pragma unroll
        for (int d = 0; d < 3; ++d)
            us[d * FV + q] = in ? node_velocity(gu, a.gnm, off, moff, d) : T(0);
    }
    __syncthreads();

    for (uint32_t s = beg + threadIdx.x; s < end; s += BT) {
        const int64_t p = int64_t(a.perm[s]);
        const int64_t xb = a.x.base(p);
        T xp[3] = {a.x.at(xb, 0), a.x.at(xb, 1), a.x.at(xb, 2)};
        Spline<ORDER, T> sd;
        eval_spline<ORDER, T, false, true>(a.lat, xp, sd);
        const int li = sd.s0[0] - f0[0], lj = sd.s0[1] - f0[1], lk = sd.s0[2] - f0[2];
        if (!(li >= 0 && li < BB && lj >= 0 && lj < BB && lk >= 0 && lk < BB)) {
            push_stale(g, s);
            continue;
        }
        const int lbase = (li * FN + lj) * FN + lk;
        T cp[12];
        apic_g2p_particle<ORDER, false>(a, sd, [&](int i, int j, int k, T* ue) {
            const int q = lbase + (i * FN + j) * FN + k;
            ue[0] = us[q];
            ue[1] = us[FV + q];
            ue[2] = us[2 * FV + q];
        }, cp);
        const int64_t cb = a.c.base(p);
#pragma unroll
        for (int r = 0; r < 12; ++r) a.c.at(cb, r) = cp[r];
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

// Cubic brick edge of the binning, or 0 when the brick kernels do not
// support it.
static int brick_edge(const pic_binning& geom)
{
    const int e = geom.brick[0];
    if (geom.brick[1] != e || geom.brick[2] != e) return 0;
    return (e == 4 || e == 8) ? e : 0;
}

template <typename T>
static bool make_brick_args(pic_context* ctx, int order, int entity, const pic_mesh& mesh,
                            int64_t np, const pic_pfield& x, const pic_pfield* m,
                            const pic_pfield& c, const pic_grid* gm, const pic_grid& gmu,
                            const pic_grid* gmass, const pic_bins& bins, BrickArgs<T>& g)
{
    const int bb = brick_edge(bins.geom);
    if (!bb || !bins.keys || !bins.bin_start) return false;
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
        g.nb[d] = (bins.geom.ncell[d] + bb - 1) / bb;
    }
    g.stale = static_cast<uint32_t*>(scratch(ctx, size_t(np + 1) * sizeof(uint32_t)));
    return g.stale != nullptr;
}

template <typename T>
struct P2GBrickLaunch {
    const BrickArgs<T>& g;
    hipStream_t st;
    template <int BB, int O, bool R>
    void go()
    {
        k_apic_p2g_brick<BB, O, T, R><<<g.nbricks, BT, 0, st>>>(g);
    }
    template <int BB>
    void by_order(int order)
    {
        constexpr bool F32 = std::is_same<T, float>::value;
        if (order == 1) go<BB, 1, true>();
        else if (order == 2) go<BB, 2, F32 ? true : PIC_F64_P2G_REGACC>();
        else go<BB, 3, false>();
    }
};

template <typename T>
struct G2PBrickLaunch {
    const BrickArgs<T>& g;
    hipStream_t st;
    template <int BB>
    void by_order(int order)
    {
        if (order == 1) k_apic_g2p_brick<BB, 1, T><<<g.nbricks, BT, 0, st>>>(g);
        else if (order == 2) k_apic_g2p_brick<BB, 2, T><<<g.nbricks, BT, 0, st>>>(g);
        else k_apic_g2p_brick<BB, 3, T><<<g.nbricks, BT, 0, st>>>(g);
    }
};

template <typename T>
static int p2g_brick_run(pic_context* ctx, int order, int entity, const pic_mesh& mesh, int64_t np,
                         const pic_pfield& x, const pic_pfield& m, const pic_pfield& c,
                         const pic_grid& gm, const pic_grid& gmu, const pic_bins& bins,
                         int dtype)
{
    BrickArgs<T> g;
    if (!make_brick_args<T>(ctx, order, entity, mesh, np, x, &m, c, &gm, gmu, nullptr, bins, g))
        return PIC_ERR_UNSUPPORTED;
    if (set_u32(ctx, g.stale, 0u)) return PIC_ERR_HIP;
    {
        TimedLaunch tl(ctx, "apic_p2g");
        P2GBrickLaunch<T> L{g, ctx->stream};
        if (brick_edge(bins.geom) == 4) L.template by_order<4>(order);
        else L.template by_order<8>(order);
        if (launch_status()) return PIC_ERR_HIP;
    }
    return apic_p2g_list(ctx, order, entity, dtype, mesh, np, x, m, c, gm, gmu, bins.perm, g.stale);
}

template <typename T>
static int g2p_brick_run(pic_context* ctx, int order, int entity, const pic_mesh& mesh, int64_t np,
                         const pic_pfield& x, const pic_grid& gu, const pic_grid* gmass,
                         const pic_pfield& c, const pic_bins& bins, int dtype)
{
    BrickArgs<T> g;
    if (!make_brick_args<T>(ctx, order, entity, mesh, np, x, nullptr, c, nullptr, gu, gmass, bins,
                            g))
        return PIC_ERR_UNSUPPORTED;
    if (set_u32(ctx, g.stale, 0u)) return PIC_ERR_HIP;
    {
        TimedLaunch tl(ctx, "apic_g2p");
        G2PBrickLaunch<T> L{g, ctx->stream};
        if (brick_edge(bins.geom) == 4) L.template by_order<4>(order);
        else L.template by_order<8>(order);
        if (launch_status()) return PIC_ERR_HIP;
    }
    return apic_g2p_list(ctx, order, entity, dtype, mesh, np, x, gu, gmass, c, bins.perm, g.stale);
}

int apic_p2g_brick(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                   int64_t np, const pic_pfield& x, const pic_pfield& m, const pic_pfield& c,
                   const pic_grid& gm, const pic_grid& gmu, const pic_bins& bins)
{
    if (dtype == PIC_F32)
        return p2g_brick_run<float>(ctx, order, entity, mesh, np, x, m, c, gm, gmu, bins, dtype);
    return p2g_brick_run<double>(ctx, order, entity, mesh, np, x, m, c, gm, gmu, bins, dtype);
}

int apic_g2p_brick(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                   int64_t np, const pic_pfield& x, const pic_grid& gu, const pic_grid* gmass,
                   const pic_pfield& c, const pic_bins& bins)
{
    if (dtype == PIC_F32)
        return g2p_brick_run<float>(ctx, order, entity, mesh, np, x, gu, gmass, c, bins, dtype);
    return g2p_brick_run<double>(ctx, order, entity, mesh, np, x, gu, gmass, c, bins, dtype);
}

}  // namespace pic
