// Generic spline transfers: the batch forms of Picasso::G2P::value /
// gradient / divergence and Picasso::P2G::value / gradient / divergence
// (Picasso_ParticleInterpolation.hpp:110-279, which forward to
// Cabana::Grid::G2P / P2G), on any entity lattice and spline order.
//
// With w_n the node weight and gw_n its physical gradient:
//   G2P value       out(p, c)      = sum_n w_n f(n, c)            c < ncomp (1..3)
//   G2P gradient    out(p, 3c + d) = sum_n gw_n[d] f(n, c)        ncomp 1 or 3
//   G2P divergence  out(p)         = sum_n sum_d gw_n[d] f(n, d)  ncomp 3
//   P2G value       f(n, c)       += w_n v(p, c)                  ncomp 1..3
//   P2G gradient    f(n, d)       += gw_n[d] v(p)                 ncomp 1
//   P2G divergence  f(n)          += sum_d gw_n[d] v(p, d)        ncomp 3
//                   f(n, i)       += sum_j gw_n[j] v(p, 3i + j)   ncomp 9
// The vector gradient's component order is pic_g2p_gradu's (du_i/dx_j at
// 3i + j).  One work-item per particle: G2P gathers the stencil, P2G adds
// into the grid with one atomic per node and component.  With bins, the
// particles are visited in sorted order, so neighbouring work-items touch
// neighbouring nodes.
#include "pic_brick.hpp"
#include "pic_common.hpp"
#include "pic_internal.hpp"

namespace pic {

template <typename T>
struct InterpArgs {
    Lattice lat;
    int64_t np;
    const uint32_t* perm;  // sorted order, or null
    PField<T> x;
    PField<T> pv;          // particle values (P2G in) or results (G2P out)
    GView<T> g;            // grid field
    int* err;
    const uint32_t* list;  // [count, slots...] to process instead of 0..np-1, or null
};

template <int ORDER, typename T>
__device__ __forceinline__ void node_weights(const Spline<ORDER, T>& sd, int i, int j, int k,
                                             T& w, T gw[3])
{
    w = sd.w[0][i] * sd.w[1][j] * sd.w[2][k];
    gw[0] = sd.g[0][i] * sd.w[1][j] * sd.w[2][k];
    gw[1] = sd.w[0][i] * sd.g[1][j] * sd.w[2][k];
    gw[2] = sd.w[0][i] * sd.w[1][j] * sd.g[2][k];
}

template <int ORDER, typename T, int OP, int NC>
__global__ __launch_bounds__(256) void k_g2p_interp(InterpArgs<T> a)
{
    constexpr int NK = Knots<ORDER>::NK;
    constexpr int NO = OP == PIC_INTERP_VALUE ? NC : OP == PIC_INTERP_GRADIENT ? 3 * NC : 1;
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= a.np) return;
    const int64_t p = a.perm ? int64_t(a.perm[s]) : s;
    const int64_t xb = a.x.base(p);
    const T xp[3] = {a.x.at(xb, 0), a.x.at(xb, 1), a.x.at(xb, 2)};
    Spline<ORDER, T> sd;
    eval_spline<ORDER, T, OP != PIC_INTERP_VALUE, false>(a.lat, xp, sd);
    if (!stencil_inside(sd, a.g)) {
        atomicOr(a.err, ERRBIT_BOUNDS);
        return;
    }
    T r[NO];
#pragma unroll
    for (int q = 0; q < NO; ++q) r[q] = T(0);
#pragma unroll
    for (int i = 0; i < NK; ++i)
#pragma unroll
        for (int j = 0; j < NK; ++j)
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                T w = T(0), gw[3] = {T(0), T(0), T(0)};
                if constexpr (OP == PIC_INTERP_VALUE) w = sd.w[0][i] * sd.w[1][j] * sd.w[2][k];
                else node_weights(sd, i, j, k, w, gw);
                const int64_t off = a.g.off(sd.s0[0] + i, sd.s0[1] + j, sd.s0[2] + k);
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    const T f = a.g.data[off + c * a.g.cs];
                    if constexpr (OP == PIC_INTERP_VALUE) {
                        r[c] += w * f;
                    } else if constexpr (OP == PIC_INTERP_GRADIENT) {
#pragma unroll
                        for (int d = 0; d < 3; ++d) r[3 * c + d] += gw[d] * f;
                    } else {
                        r[0] += gw[c] * f;
                    }
                }
            }
    const int64_t ob = a.pv.base(p);
#pragma unroll
    for (int q = 0; q < NO; ++q) a.pv.at(ob, q) = r[q];
}

template <int ORDER, typename T, int OP, int NC>
__device__ __forceinline__ void p2g_interp_particle(const InterpArgs<T>& a, int64_t p);

// Grid-stride over the particles, or over a device-side list of sorted
// slots (list[0] = count) left by the brick kernel.
template <int ORDER, typename T, int OP, int NC>
__global__ __launch_bounds__(256) void k_p2g_interp(InterpArgs<T> a)
{
    const int64_t n = a.list ? int64_t(a.list[0]) : a.np;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t s = a.list ? int64_t(a.list[1 + i]) : i;
        p2g_interp_particle<ORDER, T, OP, NC>(a, a.perm ? int64_t(a.perm[s]) : s);
    }
}

template <int ORDER, typename T, int OP, int NC>
__device__ __forceinline__ void p2g_interp_particle(const InterpArgs<T>& a, int64_t p)
{
    constexpr int NK = Knots<ORDER>::NK;
    constexpr int NG = OP == PIC_INTERP_VALUE ? NC : OP == PIC_INTERP_GRADIENT ? 3 : NC / 3;
    const int64_t xb = a.x.base(p);
    const T xp[3] = {a.x.at(xb, 0), a.x.at(xb, 1), a.x.at(xb, 2)};
    Spline<ORDER, T> sd;
    eval_spline<ORDER, T, OP != PIC_INTERP_VALUE, false>(a.lat, xp, sd);
    if (!stencil_inside(sd, a.g)) {
        atomicOr(a.err, ERRBIT_BOUNDS);
        return;
    }
    T v[NC];
    const int64_t vb = a.pv.base(p);
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = a.pv.at(vb, c);
#pragma unroll
    for (int i = 0; i < NK; ++i)
#pragma unroll
        for (int j = 0; j < NK; ++j)
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                T w = T(0), gw[3] = {T(0), T(0), T(0)};
                if constexpr (OP == PIC_INTERP_VALUE) w = sd.w[0][i] * sd.w[1][j] * sd.w[2][k];
                else node_weights(sd, i, j, k, w, gw);
                const int64_t off = a.g.off(sd.s0[0] + i, sd.s0[1] + j, sd.s0[2] + k);
#pragma unroll
                for (int c = 0; c < NG; ++c) {
                    T add;
                    if constexpr (OP == PIC_INTERP_VALUE) add = w * v[c];
                    else if constexpr (OP == PIC_INTERP_GRADIENT) add = gw[c] * v[0];
                    else add = gw[0] * v[3 * c] + gw[1] * v[3 * c + 1] + gw[2] * v[3 * c + 2];
                    atomic_add(a.g.data + off + c * a.g.cs, add);
                }
            }
}

// ---------------------------------------------------------------------------
// P2G through 4^3-bin bricks, for particles binned by pic_bin_sort on the
// spline's knot lattice (Node / Cell): the walk, staging, per-(bin,
// component) register sums, node gather and stale list of the APIC
// brick kernel (pic_brick.hpp), with this file's per-particle formulas
// factored per dimension.  Four work-items per bin, one per grid component
// (NG <= 3; the fourth idles).  Staged per particle: (f0, f1, f2, v0) and,
// as NC needs, (v1..v4), (v5..v8).
template <typename T>
struct InterpBrickArgs {
    InterpArgs<T> a;
    PIC_BRICK_WALK_FIELDS
};

constexpr int IPB = 4;  // brick edge in bins

template <int ORDER, typename T>
struct InterpBrickShape {
    static constexpr int NK = ORDER + 1;
    static constexpr int NN = NK * NK * NK;
    static constexpr int BINS = IPB * IPB * IPB;
    static constexpr int NT = BINS * 4;
    static constexpr int SLOTS = NN * BINS;
    static constexpr int SUM_BUDGET = sizeof(T) == 4 ? 34 * 1024 : 17 * 1024;
    static constexpr int NROUND = 4 * SLOTS * int(sizeof(T)) <= SUM_BUDGET   ? 1
                                  : 2 * SLOTS * int(sizeof(T)) <= SUM_BUDGET ? 2
                                                                              : 4;
    static constexpr int CPR = 4 / NROUND;
    static constexpr int CAP = sizeof(T) == 4 ? 616 : 256;
    static constexpr int STAGE = CAP * 16;
    static constexpr int SUMS = CPR * SLOTS;
    static constexpr int LDS = STAGE > SUMS ? STAGE : SUMS;
};

template <int ORDER, typename T>
using InterpStage = BrickStage<(InterpBrickShape<ORDER, T>::CAP + InterpBrickShape<ORDER, T>::NT - 1) /
                                   InterpBrickShape<ORDER, T>::NT,
                               InterpBrickShape<ORDER, T>::NT, InterpBrickShape<ORDER, T>::BINS>;

// One particle's contribution to grid component g at the bin's NK^3
// stencil nodes: acc[(i NK + j) NK + k] += the formula of OP.
template <int ORDER, typename T, int OP, int NC>
__device__ __forceinline__ void interp_brick_acc(const Spline<ORDER, T>& sd, const T (&v)[NC],
                                                 int g, T* acc)
{
    constexpr int NK = ORDER + 1;
    if constexpr (OP == PIC_INTERP_VALUE) {
        // w_i w_j w_k v_g
        const T vg = NC == 1 ? v[0] : (g == 0 ? v[0] : (g == 1 ? v[NC > 1 ? 1 : 0] : v[NC > 2 ? 2 : 0]));
#pragma unroll
        for (int i = 0; i < NK; ++i) {
            const T wi = sd.w[0][i] * vg;
#pragma unroll
            for (int j = 0; j < NK; ++j) {
                const T P = wi * sd.w[1][j];
#pragma unroll
                for (int k = 0; k < NK; ++k) acc[(i * NK + j) * NK + k] += P * sd.w[2][k];
            }
        }
    } else if constexpr (OP == PIC_INTERP_GRADIENT) {
        // (dw/dx_g) v: the gradient factor in dimension g, weights elsewhere
        T a[NK], b[NK], c[NK];
#pragma unroll
        for (int n = 0; n < NK; ++n) {
            a[n] = (g == 0 ? sd.g[0][n] : sd.w[0][n]) * v[0];
            b[n] = g == 1 ? sd.g[1][n] : sd.w[1][n];
            c[n] = g == 2 ? sd.g[2][n] : sd.w[2][n];
        }
#pragma unroll
        for (int i = 0; i < NK; ++i)
#pragma unroll
            for (int j = 0; j < NK; ++j) {
                const T P = a[i] * b[j];
#pragma unroll
                for (int k = 0; k < NK; ++k) acc[(i * NK + j) * NK + k] += P * c[k];
            }
    } else {
        // grad w . (u0, u1, u2) with u = v (vector) or row g of v (tensor)
        T u0 = v[0], u1 = v[NC > 1 ? 1 : 0], u2 = v[NC > 2 ? 2 : 0];
        if constexpr (NC == 9) {
            u0 = g == 0 ? v[0] : (g == 1 ? v[3] : v[6]);
            u1 = g == 0 ? v[1] : (g == 1 ? v[4] : v[7]);
            u2 = g == 0 ? v[2] : (g == 1 ? v[5] : v[8]);
        }
#pragma unroll
        for (int i = 0; i < NK; ++i) {
            const T gi = sd.g[0][i] * u0, wi1 = sd.w[0][i] * u1, wi2 = sd.w[0][i] * u2;
#pragma unroll
            for (int j = 0; j < NK; ++j) {
                const T al = gi * sd.w[1][j] + wi1 * sd.g[1][j];
                const T be = wi2 * sd.w[1][j];
#pragma unroll
                for (int k = 0; k < NK; ++k)
                    acc[(i * NK + j) * NK + k] += al * sd.w[2][k] + be * sd.g[2][k];
            }
        }
    }
}

template <int ORDER, typename T, int OP, int NC>
__device__ __forceinline__ void interp_p2g_one_brick(const InterpBrickArgs<T>& g,
                                                     const BrickItem& it,
                                                     BrickWalk<true, InterpBrickArgs<T>>& w,
                                                     T* lds, InterpStage<ORDER, T>& st)
{
    using S = InterpBrickShape<ORDER, T>;
    using T4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    constexpr int NK = S::NK, NN = S::NN, BINS = S::BINS, NT = S::NT, CAP = S::CAP;
    constexpr int K = (CAP + NT - 1) / NT;
    constexpr int NG = OP == PIC_INTERP_VALUE ? NC : OP == PIC_INTERP_GRADIENT ? 3 : NC / 3;
    constexpr int NGRP = 1 + (NC + 2) / 4;  // staged groups: (f, v0), (v1..v4), (v5..v8)
    const InterpArgs<T>& a = g.a;
    const uint32_t beg = it.beg, end = it.end;
    int f0[3];
    brick_origin<IPB>(g, it.b, f0);
    const uint32_t kb = it.b * BINS;
    const int l = int(threadIdx.x) >> 2, comp = int(threadIdx.x) & 3;
    T acc[NN];
#pragma unroll
    for (int n = 0; n < NN; ++n) acc[n] = T(0);
    for (uint32_t c0 = beg; c0 < end; c0 += CAP) {
        const uint32_t c1 = min(end, c0 + uint32_t(CAP));
        uint32_t* rng = st.begin_chunk(g, a.perm, beg, c0, c1);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t s = c0 + threadIdx.x + uint32_t(k * NT);
            if (s >= c1) continue;
            const int64_t p = int64_t(st.cs.perm[k]);
            const int64_t xb = a.x.base(p), vb = a.pv.base(p);
            const int q = int(s - c0);
            const int lb = int(st.cs.key[k] - kb);
            const int eb[3] = {f0[0] + lb / (IPB * IPB), f0[1] + (lb / IPB) % IPB,
                               f0[2] + lb % IPB};
            T vals[12];
#pragma unroll
            for (int c = 0; c < 12; ++c) vals[c] = c < NC ? a.pv.at(vb, c) : T(0);
            T4 v4;
            T* fr = &v4.x;
            bool ok = true;
#pragma unroll
            for (int dim = 0; dim < 3; ++dim) {
                double fd;
                const double kb0 = knot_base<ORDER>(double(a.x.at(xb, dim)), a.lat.o[dim],
                                                     a.lat.rdx, fd);
                ok = ok && int(kb0) + Knots<ORDER>::FIRST == eb[dim];
                fr[dim] = T(fd);
            }
            if (!ok) {
                push_stale(g, s);
                v4.x = T(STALE_F);
            }
            v4.w = vals[0];
            *reinterpret_cast<T4*>(lds + stage_at(q, 0)) = v4;
#pragma unroll
            for (int gg = 1; gg < NGRP; ++gg) {
                v4.x = vals[4 * gg - 3];
                v4.y = vals[4 * gg - 2];
                v4.z = vals[4 * gg - 1];
                v4.w = vals[4 * gg];
                *reinterpret_cast<T4*>(lds + stage_at(q, gg)) = v4;
            }
            st.cs.template note<BINS>(k, q, kb, rng);
        }
        st.end_chunk();
        __syncthreads();
        if (c0 == beg) w.fetch();
        const uint32_t q0 = rng[l], q1 = rng[BINS + l];
        if (comp < NG) {
            for (uint32_t q = q0; q < q1; ++q) {
                const T4 fv = *reinterpret_cast<const T4*>(lds + stage_at(int(q), 0));
                if (fv.x == T(STALE_F)) continue;
                T v[NC];
                v[0] = fv.w;
#pragma unroll
                for (int gg = 1; gg < NGRP; ++gg) {
                    const T4 u = *reinterpret_cast<const T4*>(lds + stage_at(int(q), gg));
                    const T uu[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (4 * gg - 3 + e < NC) v[4 * gg - 3 + e] = uu[e];
                }
                const T fr[3] = {fv.x, fv.y, fv.z};
                Spline<ORDER, T> sd;
                spline_from_frac<ORDER, T, OP != PIC_INTERP_VALUE, false>(a.lat, fr, sd);
                interp_brick_acc<ORDER, T, OP, NC>(sd, v, comp, acc);
            }
        }
        __syncthreads();
    }
    for (int r = 0; r < S::NROUND; ++r) {
        const int r_lo = r * S::CPR;
        if (comp >= r_lo && comp < r_lo + S::CPR) {
            T* dst = lds + l * S::CPR + (comp - r_lo);
#pragma unroll
            for (int n = 0; n < NN; ++n) dst[n * BINS * S::CPR] = acc[n];
        }
        if (r == 0) w.publish();
        __syncthreads();
        if (r == 0) st.prefetch_next(g, a.perm, w, CAP);
        gather_flush<NK, IPB, NT, S::CPR>(lds, f0, r_lo, 0, GView<T>{}, a.g, a.err);
        if (r + 1 < S::NROUND) __syncthreads();
    }
}

template <int ORDER, typename T, int OP, int NC>
__global__ __launch_bounds__((InterpBrickShape<ORDER, T>::NT)) void k_p2g_interp_brick(
    InterpBrickArgs<T> g)
{
    using S = InterpBrickShape<ORDER, T>;
    __shared__ __align__(16) T lds[S::LDS];
    __shared__ uint32_t rng[2 * 2 * S::BINS];
    __shared__ BrickItem q[2];
    for (int i = threadIdx.x; i < 4 * S::BINS; i += S::NT) rng[i] = 0u;
    InterpStage<ORDER, T> st{rng, 0, false, {}};
    walk_bricks<true>(g, q, [&](const BrickItem& it, BrickWalk<true, InterpBrickArgs<T>>& w) {
        interp_p2g_one_brick<ORDER, T, OP, NC>(g, it, w, lds, st);
    });
}

// Components of the particle side and of the grid side of (op, ncomp), or
// false for an unsupported pair.  ncomp counts the input's components.
static bool interp_shape(bool g2p, int op, int ncomp, int* nparticle, int* ngrid)
{
    if (g2p) {
        *ngrid = ncomp;
        if (op == PIC_INTERP_VALUE && ncomp >= 1 && ncomp <= 3) *nparticle = ncomp;
        else if (op == PIC_INTERP_GRADIENT && (ncomp == 1 || ncomp == 3)) *nparticle = 3 * ncomp;
        else if (op == PIC_INTERP_DIVERGENCE && ncomp == 3) *nparticle = 1;
        else return false;
    } else {
        *nparticle = ncomp;
        if (op == PIC_INTERP_VALUE && ncomp >= 1 && ncomp <= 3) *ngrid = ncomp;
        else if (op == PIC_INTERP_GRADIENT && ncomp == 1) *ngrid = 3;
        else if (op == PIC_INTERP_DIVERGENCE && (ncomp == 3 || ncomp == 9)) *ngrid = ncomp / 3;
        else return false;
    }
    return true;
}

template <typename T, bool G2P>
struct InterpLaunch {
    const InterpArgs<T>& a;
    unsigned blocks;
    hipStream_t st;
    template <int O, int OP, int NC>
    void go()
    {
        if constexpr (G2P) k_g2p_interp<O, T, OP, NC><<<blocks, 256, 0, st>>>(a);
        else k_p2g_interp<O, T, OP, NC><<<blocks, 256, 0, st>>>(a);
    }
    template <int O>
    void by_shape(int op, int nc)
    {
        if (op == PIC_INTERP_VALUE) {
            if (nc == 1) go<O, PIC_INTERP_VALUE, 1>();
            else if (nc == 2) go<O, PIC_INTERP_VALUE, 2>();
            else go<O, PIC_INTERP_VALUE, 3>();
        } else if (op == PIC_INTERP_GRADIENT) {
            if (nc == 1) go<O, PIC_INTERP_GRADIENT, 1>();
            else if constexpr (G2P) go<O, PIC_INTERP_GRADIENT, 3>();
        } else {
            if (nc == 3) go<O, PIC_INTERP_DIVERGENCE, 3>();
            else if constexpr (!G2P) go<O, PIC_INTERP_DIVERGENCE, 9>();
        }
    }
    void run(int order, int op, int nc)
    {
        if (order == 1) by_shape<1>(op, nc);
        else if (order == 2) by_shape<2>(op, nc);
        else by_shape<3>(op, nc);
    }
};

template <typename T>
struct InterpBrickLaunch {
    pic_context* ctx;
    const InterpBrickArgs<T>& g;
    template <int O, int OP, int NC>
    void go()
    {
        constexpr int NT = InterpBrickShape<O, T>::NT;
        const unsigned nbk =
            persistent_blocks(ctx, k_p2g_interp_brick<O, T, OP, NC>, NT, 0, g.nbricks);
        k_p2g_interp_brick<O, T, OP, NC><<<nbk, NT, 0, ctx->stream>>>(g);
    }
    template <int O>
    void by_shape(int op, int nc)
    {
        if (op == PIC_INTERP_VALUE) {
            if (nc == 1) go<O, PIC_INTERP_VALUE, 1>();
            else if (nc == 2) go<O, PIC_INTERP_VALUE, 2>();
            else go<O, PIC_INTERP_VALUE, 3>();
        } else if (op == PIC_INTERP_GRADIENT) {
            go<O, PIC_INTERP_GRADIENT, 1>();
        } else {
            if (nc == 3) go<O, PIC_INTERP_DIVERGENCE, 3>();
            else go<O, PIC_INTERP_DIVERGENCE, 9>();
        }
    }
    void run(int order, int op, int nc)
    {
        if (order == 1) by_shape<1>(op, nc);
        else if (order == 2) by_shape<2>(op, nc);
        else by_shape<3>(op, nc);
    }
};

// Binned P2G on the brick kernel, then the particles it left in the stale
// list through the per-particle kernel; PIC_ERR_UNSUPPORTED when the bins
// do not fit the brick kernel (the caller then runs the per-particle one).
template <typename T>
static int p2g_interp_brick_run(pic_context* ctx, int op, int order, int entity, int ncomp,
                                const InterpArgs<T>& a, const pic_bins& bins)
{
    if (brick_edge(bins.geom) != IPB) return PIC_ERR_UNSUPPORTED;
    InterpBrickArgs<T> g;
    if (const int rc = setup_walk(ctx, order, entity, a.np, bins, g)) return rc;
    g.a = a;
    {
        TimedLaunch tl(ctx, "p2g_interp");
        InterpBrickLaunch<T>{ctx, g}.run(order, op, ncomp);
        if (launch_status()) return PIC_ERR_HIP;
    }
    InterpArgs<T> la = g.a;
    la.list = g.stale;
    TimedLaunch tl(ctx, "p2g_interp_stale");
    InterpLaunch<T, false>{la, 512u, ctx->stream}.run(order, op, ncomp);
    return launch_status();
}

static int interp_run(pic_context* ctx, bool g2p, int op, int order, int entity, int dtype,
                      const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                      const pic_pfield* pv, int ncomp, const pic_grid* field, uint32_t flags,
                      const pic_bins* bins)
{
    int npart = 0, ngrid = 0;
    if (!ctx || !mesh || !(mesh->h > 0.0) || order < 1 || order > 3 || !valid_entity(entity) ||
        (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 || np > int64_t(UINT32_MAX) ||
        !valid_grid(field) || !interp_shape(g2p, op, ncomp, &npart, &ngrid) ||
        (flags & ~uint32_t(PIC_P2G_ZERO_FIRST)) != 0 || (g2p && flags != 0))
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !valid_pfield(pv))) return PIC_ERR_ARG;
    if (bins && !bins->perm) return PIC_ERR_ARG;
    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;
    if (flags & PIC_P2G_ZERO_FIRST) {
        const int rc = zero_grid(ctx, dtype, *field, ngrid);
        if (rc != PIC_OK) return rc;
    }
    if (np == 0) return launch_status();
    const unsigned b = blocks_for(np, 256);
    const uint32_t* perm = bins ? bins->perm : nullptr;
    auto go = [&](auto tag) {
        using T = decltype(tag);
        const InterpArgs<T> a{make_lattice(*mesh, entity), np, perm, make_pfield<T>(*x),
                              make_pfield<T>(*pv), make_gview<T>(*field), ctx->d_err, nullptr};
        if (!g2p && bins) {
            const int rc = p2g_interp_brick_run<T>(ctx, op, order, entity, ncomp, a, *bins);
            if (rc != PIC_ERR_UNSUPPORTED) return rc;
        }
        TimedLaunch tl(ctx, g2p ? "g2p_interp" : "p2g_interp");
        if (g2p) InterpLaunch<T, true>{a, b, ctx->stream}.run(order, op, ncomp);
        else InterpLaunch<T, false>{a, b, ctx->stream}.run(order, op, ncomp);
        return launch_status();
    };
    return dtype == PIC_F64 ? go(double{}) : go(float{});
}

}  // namespace pic

using namespace pic;

extern "C" int pic_g2p_interp(pic_context* ctx, int op, int order, int entity, int dtype,
                              const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                              const pic_grid* field, int ncomp, const pic_pfield* out,
                              const pic_bins* bins)
{
    Region region("picasso_amd::G2P");
    return interp_run(ctx, true, op, order, entity, dtype, mesh, np, x, out, ncomp, field, 0u,
                      bins);
}

extern "C" int pic_p2g_interp(pic_context* ctx, int op, int order, int entity, int dtype,
                              const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                              const pic_pfield* value, int ncomp, const pic_grid* field,
                              uint32_t flags, const pic_bins* bins)
{
    Region region("picasso_amd::P2G");
    return interp_run(ctx, false, op, order, entity, dtype, mesh, np, x, value, ncomp, field,
                      flags, bins);
}
