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
__global__ __launch_bounds__(256) void k_p2g_interp(InterpArgs<T> a)
{
    constexpr int NK = Knots<ORDER>::NK;
    constexpr int NG = OP == PIC_INTERP_VALUE ? NC : OP == PIC_INTERP_GRADIENT ? 3 : NC / 3;
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
    TimedLaunch tl(ctx, g2p ? "g2p_interp" : "p2g_interp");
    const unsigned b = blocks_for(np, 256);
    const uint32_t* perm = bins ? bins->perm : nullptr;
    if (dtype == PIC_F64) {
        InterpArgs<double> a{make_lattice(*mesh, entity), np, perm, make_pfield<double>(*x),
                             make_pfield<double>(*pv), make_gview<double>(*field), ctx->d_err};
        if (g2p) InterpLaunch<double, true>{a, b, ctx->stream}.run(order, op, ncomp);
        else InterpLaunch<double, false>{a, b, ctx->stream}.run(order, op, ncomp);
    } else {
        InterpArgs<float> a{make_lattice(*mesh, entity), np, perm, make_pfield<float>(*x),
                            make_pfield<float>(*pv), make_gview<float>(*field), ctx->d_err};
        if (g2p) InterpLaunch<float, true>{a, b, ctx->stream}.run(order, op, ncomp);
        else InterpLaunch<float, false>{a, b, ctx->stream}.run(order, op, ncomp);
    }
    return launch_status();
}

}  // namespace pic

using namespace pic;

extern "C" int pic_g2p_interp(pic_context* ctx, int op, int order, int entity, int dtype,
                              const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                              const pic_grid* field, int ncomp, const pic_pfield* out,
                              const pic_bins* bins)
{
    return interp_run(ctx, true, op, order, entity, dtype, mesh, np, x, out, ncomp, field, 0u,
                      bins);
}

extern "C" int pic_p2g_interp(pic_context* ctx, int op, int order, int entity, int dtype,
                              const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                              const pic_pfield* value, int ncomp, const pic_grid* field,
                              uint32_t flags, const pic_bins* bins)
{
    return interp_run(ctx, false, op, order, entity, dtype, mesh, np, x, value, ncomp, field,
                      flags, bins);
}
