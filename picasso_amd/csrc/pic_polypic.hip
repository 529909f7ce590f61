// PolyPIC particle<->grid transfers (Picasso_PolyPIC.hpp) for gfx950.
//
// The reference defines the linear spline with 8 modes per component
// (basis 1, y0, y1, y2, y0y1, y0y2, y1y2, y0y1y2).  Orders 2 and 3 use 27
// modes, the basis {1, y, y^2}^3: the first eight modes are the linear ones
// in the reference's order, followed by the 19 modes holding a square in
// lexicographic order of their exponents.  g2p mode coefficients are the
// Taylor coefficients of the grid interpolant at the particle: per dimension
// the weight, its gradient, or half its second derivative.
#include "pic_common.hpp"
#include "pic_internal.hpp"

namespace pic {

__device__ constexpr int MODE_EXP[27][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1},
    {1, 1, 1}, {0, 0, 2}, {0, 1, 2}, {0, 2, 0}, {0, 2, 1}, {0, 2, 2}, {1, 0, 2},
    {1, 1, 2}, {1, 2, 0}, {1, 2, 1}, {1, 2, 2}, {2, 0, 0}, {2, 0, 1}, {2, 0, 2},
    {2, 1, 0}, {2, 1, 1}, {2, 1, 2}, {2, 2, 0}, {2, 2, 1}, {2, 2, 2}};

template <int ORDER>
struct Modes {
    static constexpr int NM = ORDER == 1 ? 8 : 27;
};

template <typename T>
struct PolyArgs {
    Lattice lat;
    int64_t np;
    const uint32_t* perm;
    PField<T> x, m, u, c;  // positions, mass, velocity modes (NM x 3), field modes
    GView<T> gmu, gm;
    int pcols;             // columns of the particle field (ncomp)
    int col;               // particle column a staggered transfer works on
    double dt;
    int* err;
};

template <typename T>
__device__ __forceinline__ T ypow(T y, int e)
{
    return e == 0 ? T(1) : (e == 1 ? y : y * y);
}

// PolyPIC::p2g (Picasso_PolyPIC.hpp:39-130, 138-229).  NC is the number of
// grid components (1 for staggered).
template <int ORDER, int NC, bool STAG, typename T>
__global__ __launch_bounds__(256) void k_polypic_p2g(PolyArgs<T> a)
{
    constexpr int NK = Knots<ORDER>::NK;
    constexpr int NM = Modes<ORDER>::NM;
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= a.np) return;
    const int64_t p = a.perm ? int64_t(a.perm[s]) : s;

    const int64_t xb = a.x.base(p);
    T xp[3] = {a.x.at(xb, 0), a.x.at(xb, 1), a.x.at(xb, 2)};
    const T mp = a.m.at(a.m.base(p), 0);
    Spline<ORDER, T> sd;
    eval_spline<ORDER, T, false, true>(a.lat, xp, sd);
    if (!stencil_inside(sd, a.gmu)) {
        atomicOr(a.err, ERRBIT_BOUNDS);
        return;
    }

    // Affine material motion operator and its inverse (:81-87).
    const int64_t ub = a.u.base(p);
    const T dt = (T)a.dt;
    T am[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int d = 0; d < 3; ++d)
            am[r * 3 + d] = (r == d ? T(1) : T(0)) + dt * a.u.at(ub, (1 + r) * 3 + d);
    T ai[9];
    inverse3(am, det3(am), ai);

    const int64_t cb = a.c.base(p);
    T cm[NM][NC];
#pragma unroll
    for (int r = 0; r < NM; ++r)
#pragma unroll
        for (int d = 0; d < NC; ++d)
            cm[r][d] = STAG ? a.c.at(cb, r * a.pcols + a.col) : a.c.at(cb, r * NC + d);

#pragma unroll
    for (int i = 0; i < NK; ++i)
#pragma unroll
        for (int j = 0; j < NK; ++j)
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                const T wm = mp * sd.w[0][i] * sd.w[1][j] * sd.w[2][k];
                const T dist[3] = {sd.d[0][i], sd.d[1][j], sd.d[2][k]};
                T y[3];
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    y[q] = ai[q * 3 + 0] * dist[0] + ai[q * 3 + 1] * dist[1] + ai[q * 3 + 2] * dist[2];
                T basis[NM];
#pragma unroll
                for (int r = 0; r < NM; ++r)
                    basis[r] = ypow(y[0], MODE_EXP[r][0]) * ypow(y[1], MODE_EXP[r][1]) *
                               ypow(y[2], MODE_EXP[r][2]);
                const int64_t off = a.gmu.off(sd.s0[0] + i, sd.s0[1] + j, sd.s0[2] + k);
#pragma unroll
                for (int d = 0; d < NC; ++d) {
                    // wm * ~c_p.column(d) * basis
                    T v = T(0);
#pragma unroll
                    for (int r = 0; r < NM; ++r) v += (wm * cm[r][d]) * basis[r];
                    atomic_add(a.gmu.data + off + d * a.gmu.cs, v);
                }
                atomic_add(a.gm.data + a.gm.off(sd.s0[0] + i, sd.s0[1] + j, sd.s0[2] + k), wm);
            }
}

// PolyPIC::g2p (Picasso_PolyPIC.hpp:235-293, 299-359).
template <int ORDER, int NC, bool STAG, typename T>
__global__ __launch_bounds__(256) void k_polypic_g2p(PolyArgs<T> a)
{
    constexpr int NK = Knots<ORDER>::NK;
    constexpr int NM = Modes<ORDER>::NM;
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= a.np) return;
    const int64_t p = a.perm ? int64_t(a.perm[s]) : s;
    const int64_t xb = a.x.base(p);
    T xp[3] = {a.x.at(xb, 0), a.x.at(xb, 1), a.x.at(xb, 2)};
    Spline<ORDER, T> sd;
    eval_spline<ORDER, T, true, false, (NM > 8)>(a.lat, xp, sd);
    if (!stencil_inside(sd, a.gmu)) {
        atomicOr(a.err, ERRBIT_BOUNDS);
        return;
    }
    T cm[NM][NC];
#pragma unroll
    for (int r = 0; r < NM; ++r)
#pragma unroll
        for (int d = 0; d < NC; ++d) cm[r][d] = T(0);
#pragma unroll
    for (int i = 0; i < NK; ++i)
#pragma unroll
        for (int j = 0; j < NK; ++j)
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                T fac[3][3];
                const int idx[3] = {i, j, k};
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    fac[q][0] = sd.w[q][idx[q]];
                    fac[q][1] = sd.g[q][idx[q]];
                    if constexpr (NM > 8) fac[q][2] = T(0.5) * sd.g2[q][idx[q]];
                    else fac[q][2] = T(0);
                }
                const int64_t off = a.gmu.off(sd.s0[0] + i, sd.s0[1] + j, sd.s0[2] + k);
                T ue[NC];
#pragma unroll
                for (int d = 0; d < NC; ++d) ue[d] = a.gmu.data[off + d * a.gmu.cs];
#pragma unroll
                for (int r = 0; r < NM; ++r) {
                    const T coeff = fac[0][MODE_EXP[r][0]] * fac[1][MODE_EXP[r][1]] *
                                    fac[2][MODE_EXP[r][2]];
#pragma unroll
                    for (int d = 0; d < NC; ++d) cm[r][d] += coeff * ue[d];
                }
            }
    const int64_t cb = a.c.base(p);
    if constexpr (STAG) {
        // Only the working column is reset and rewritten (:328).
#pragma unroll
        for (int r = 0; r < NM; ++r) a.c.at(cb, r * a.pcols + a.col) = cm[r][0];
    } else {
#pragma unroll
        for (int r = 0; r < NM; ++r)
#pragma unroll
            for (int d = 0; d < NC; ++d) a.c.at(cb, r * NC + d) = cm[r][d];
    }
}

template <typename T>
struct PolyLaunch {
    PolyArgs<T> a;
    hipStream_t st;
    bool p2g;
    template <int O, int NC, bool S>
    void go()
    {
        const unsigned b = blocks_for(a.np, 256);
        if (p2g) k_polypic_p2g<O, NC, S, T><<<b, 256, 0, st>>>(a);
        else k_polypic_g2p<O, NC, S, T><<<b, 256, 0, st>>>(a);
    }
    template <int O>
    void by_nc(int nc, bool stag)
    {
        if (stag) go<O, 1, true>();
        else if (nc == 1) go<O, 1, false>();
        else if (nc == 2) go<O, 2, false>();
        else go<O, 3, false>();
    }
    void run(int order, int nc, bool stag)
    {
        if (order == 1) by_nc<1>(nc, stag);
        else if (order == 2) by_nc<2>(nc, stag);
        else by_nc<3>(nc, stag);
    }
};

template <typename T>
static PolyArgs<T> make_poly_args(pic_context* ctx, int entity, int ncomp, const pic_mesh& mesh,
                                  int64_t np, const pic_pfield* x, const pic_pfield* m,
                                  const pic_pfield* u, const pic_pfield* c, const pic_grid& gmu,
                                  const pic_grid* gm, double dt, const uint32_t* perm)
{
    PolyArgs<T> a;
    a.lat = make_lattice(mesh, entity);
    a.np = np;
    a.perm = perm;
    a.x = make_pfield<T>(*x);
    a.m = m ? make_pfield<T>(*m) : PField<T>{};
    a.u = u ? make_pfield<T>(*u) : PField<T>{};
    a.c = make_pfield<T>(*c);
    a.gmu = make_gview<T>(gmu);
    a.gm = gm ? make_gview<T>(*gm) : GView<T>{};
    const int sd = stag_dim(entity);
    // Staggered: the column is entity dim for a 3-column particle field,
    // else 0 (:181, :325).
    a.pcols = ncomp;
    a.col = (sd >= 0 && ncomp == 3) ? sd : 0;
    a.dt = dt;
    a.err = ctx->d_err;
    return a;
}

}  // namespace pic

using namespace pic;

extern "C" int pic_polypic_nmode(int order)
{
    if (order < 1 || order > 3) return PIC_ERR_ARG;
    return order == 1 ? 8 : 27;
}

extern "C" int pic_polypic_p2g(pic_context* ctx, int order, int entity, int dtype,
                               const pic_mesh* mesh, int64_t np, int ncomp,
                               const pic_pfield* x, const pic_pfield* m_p,
                               const pic_pfield* u_p, const pic_pfield* c_p,
                               const pic_grid* mu_i, const pic_grid* m_i, double dt,
                               const pic_bins* bins, uint32_t flags)
{
    if (!ctx || !mesh || !(mesh->h > 0.0) || order < 1 || order > 3 ||
        !valid_entity(entity) || (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 ||
        np > int64_t(UINT32_MAX) || ncomp < 1 || ncomp > 3 || !valid_grid(mu_i) ||
        !valid_grid(m_i))
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !valid_pfield(m_p) || !valid_pfield(u_p) ||
                   !valid_pfield(c_p)))
        return PIC_ERR_ARG;
    if (bins && !bins->perm) return PIC_ERR_ARG;
    for (int d = 0; d < 3; ++d)
        if (mu_i->lo[d] != m_i->lo[d] || mu_i->n[d] != m_i->n[d]) return PIC_ERR_ARG;
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    const bool stag = stag_dim(entity) >= 0;
    if (flags & PIC_P2G_ZERO_FIRST) {
        int rc = zero_grid(ctx, dtype, *m_i, 1);
        if (rc) return rc;
        rc = zero_grid(ctx, dtype, *mu_i, stag ? 1 : ncomp);
        if (rc) return rc;
    }
    if (np == 0) return PIC_OK;
    const uint32_t* perm = bins ? bins->perm : nullptr;
    TimedLaunch tl(ctx, "polypic_p2g");
    if (dtype == PIC_F64) {
        PolyLaunch<double> L{make_poly_args<double>(ctx, entity, ncomp, *mesh, np, x, m_p, u_p, c_p,
                                                    *mu_i, m_i, dt, perm),
                             ctx->stream, true};
        L.run(order, ncomp, stag);
    } else {
        PolyLaunch<float> L{make_poly_args<float>(ctx, entity, ncomp, *mesh, np, x, m_p, u_p, c_p,
                                                  *mu_i, m_i, dt, perm),
                            ctx->stream, true};
        L.run(order, ncomp, stag);
    }
    return launch_status();
}

extern "C" int pic_polypic_g2p(pic_context* ctx, int order, int entity, int dtype,
                               const pic_mesh* mesh, int64_t np, int ncomp,
                               const pic_pfield* x, const pic_grid* u_i,
                               const pic_pfield* c_p, const pic_bins* bins)
{
    if (!ctx || !mesh || !(mesh->h > 0.0) || order < 1 || order > 3 ||
        !valid_entity(entity) || (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 ||
        np > int64_t(UINT32_MAX) || ncomp < 1 || ncomp > 3 || !valid_grid(u_i))
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !valid_pfield(c_p))) return PIC_ERR_ARG;
    if (bins && !bins->perm) return PIC_ERR_ARG;
    if (np == 0) return PIC_OK;
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    const bool stag = stag_dim(entity) >= 0;
    const uint32_t* perm = bins ? bins->perm : nullptr;
    TimedLaunch tl(ctx, "polypic_g2p");
    if (dtype == PIC_F64) {
        PolyLaunch<double> L{make_poly_args<double>(ctx, entity, ncomp, *mesh, np, x, nullptr,
                                                    nullptr, c_p, *u_i, nullptr, 0.0, perm),
                             ctx->stream, false};
        L.run(order, ncomp, stag);
    } else {
        PolyLaunch<float> L{make_poly_args<float>(ctx, entity, ncomp, *mesh, np, x, nullptr,
                                                  nullptr, c_p, *u_i, nullptr, 0.0, perm),
                            ctx->stream, false};
        L.run(order, ncomp, stag);
    }
    return launch_status();
}
