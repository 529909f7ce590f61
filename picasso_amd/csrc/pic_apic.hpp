// Per-particle APIC arithmetic shared by the scatter and brick kernels.
// Restates Picasso_APIC.hpp: p2g orders 2/3 (:55-140 collocated, :148-232
// staggered), p2g order 1 (:240-319, :327-402), g2p (:407-472, :477-544).
#pragma once

#include "pic_common.hpp"
#include "pic_internal.hpp"

namespace pic {

template <typename T>
struct ApicArgs {
    Lattice lat;
    int64_t np;
    const uint32_t* perm;  // sorted slot -> particle, or null (identity)
    const uint32_t* list;  // [count, slots...] to process instead of 0..np-1, or null
    // p2g over a list: when the count is at most serial_max (non-zero), one
    // work-item applies the particles one after another in list order.
    uint32_t serial_max;
    PField<T> x, m, c;     // positions (3), mass (1), c_p (12)
    GView<T> gm, gmu;      // grid mass, grid momentum (or velocity for g2p)
    GView<T> gnm;          // g2p with fused normalisation: grid mass (else null data)
    int sdim;              // staggered component (entity dim)
    int* err;
    // g2p with the deformation-gradient update fused in (F.data null: off).
    PField<double> F, det, Finv;
    double dt;
};

// Deformation-gradient outputs of a fused g2p (pic_apic_g2p_update_F).
struct FUpdate {
    const pic_pfield* F;
    const pic_pfield* det;
    const pic_pfield* Finv;
    double dt;
};

// Velocity gradient L of particle p applied to its F: F <- (I + dt L) F,
// det and F^{-1} stored (FP64).
template <typename T>
__device__ __forceinline__ void store_F_update(const ApicArgs<T>& a, int64_t p, const T L[9])
{
    const int64_t fb = a.F.base(p), ib = a.Finv.base(p);
    double Ld[9], Fp[9], inv[9], dd;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        Ld[q] = double(L[q]);
        Fp[q] = a.F.at(fb, q);
    }
    update_F(Ld, a.dt, Fp, dd, inv);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        a.F.at(fb, q) = Fp[q];
        a.Finv.at(ib, q) = inv[q];
    }
    a.det.at(a.det.base(p), 0) = dd;
}

// Particle state for p2g.  Staggered transfers keep only row `sdim` of
// u_p and B_p, stored at index 0.
template <typename T>
struct ApicParticle {
    T x[3];
    T m;
    T c0[3];
    T B[3][3];
};

template <bool STAG, typename T>
__device__ __forceinline__ void load_apic_particle(const ApicArgs<T>& a, int64_t p,
                                                   ApicParticle<T>& pt)
{
    const int64_t xb = a.x.base(p);
#pragma unroll
    for (int d = 0; d < 3; ++d) pt.x[d] = a.x.at(xb, d);
    pt.m = a.m.at(a.m.base(p), 0);
    const int64_t cb = a.c.base(p);
    // B_p is sliced from c_p and transposed: B(d, a) = c_p(1 + a, d)
    // (Picasso_APIC.hpp:94-101).
    if constexpr (STAG) {
        const int d = a.sdim;
        pt.c0[0] = a.c.at(cb, d);
#pragma unroll
        for (int q = 0; q < 3; ++q) pt.B[0][q] = a.c.at(cb, (1 + q) * 3 + d);
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            pt.c0[d] = a.c.at(cb, d);
#pragma unroll
            for (int q = 0; q < 3; ++q) pt.B[d][q] = a.c.at(cb, (1 + q) * 3 + d);
        }
    }
}

// Contributions of one particle to its stencil nodes, handed to
// emit(i, j, k, mu[ncomp], m) in the reference's i-outer, k-inner order.
template <int ORDER, bool STAG, typename T, typename Emit>
__device__ __forceinline__ void apic_p2g_particle(const ApicArgs<T>& a,
                                                  const ApicParticle<T>& pt,
                                                  const Spline<ORDER, T>& sd, Emit&& emit)
{
    constexpr int NK = Knots<ORDER>::NK;
    constexpr int NC = STAG ? 1 : 3;
    if constexpr (ORDER >= 2) {
        // APIC::inertialScaling (Picasso_APIC.hpp:33-47).
        const T hh = (T)a.lat.h;
        const T Dinv = T(ORDER == 2 ? 4.0 : 3.0) / (hh * hh);
        T DB[NC][3];
#pragma unroll
        for (int d = 0; d < NC; ++d)
#pragma unroll
            for (int q = 0; q < 3; ++q) DB[d][q] = Dinv * pt.B[d][q];
        // v_n = u_p + D^{-1} B_p d_n, built up one dimension at a time so
        // each node costs one multiply-add per component.
#pragma unroll
        for (int i = 0; i < NK; ++i) {
            T vi[NC];
#pragma unroll
            for (int d = 0; d < NC; ++d) vi[d] = pt.c0[d] + DB[d][0] * sd.d[0][i];
            const T wim = sd.w[0][i] * pt.m;
#pragma unroll
            for (int j = 0; j < NK; ++j) {
                T vij[NC];
#pragma unroll
                for (int d = 0; d < NC; ++d) vij[d] = vi[d] + DB[d][1] * sd.d[1][j];
                const T wijm = wim * sd.w[1][j];
#pragma unroll
                for (int k = 0; k < NK; ++k) {
                    const T wm = wijm * sd.w[2][k];
                    T mu[NC];
#pragma unroll
                    for (int d = 0; d < NC; ++d) mu[d] = wm * (vij[d] + DB[d][2] * sd.d[2][k]);
                    emit(i, j, k, mu, wm);
                }
            }
        }
    } else {
        // Gradient form, no inertial scaling (Picasso_APIC.hpp:287-318).
#pragma unroll
        for (int i = 0; i < NK; ++i)
#pragma unroll
            for (int j = 0; j < NK; ++j)
#pragma unroll
                for (int k = 0; k < NK; ++k) {
                    const T wm = sd.w[0][i] * sd.w[1][j] * sd.w[2][k] * pt.m;
                    const T gm[3] = {sd.g[0][i] * sd.w[1][j] * sd.w[2][k] * pt.m,
                                     sd.w[0][i] * sd.g[1][j] * sd.w[2][k] * pt.m,
                                     sd.w[0][i] * sd.w[1][j] * sd.g[2][k] * pt.m};
                    T mu[NC];
#pragma unroll
                    for (int d = 0; d < NC; ++d)
                        mu[d] = wm * pt.c0[d] +
                                (pt.B[d][0] * gm[0] + pt.B[d][1] * gm[1] + pt.B[d][2] * gm[2]);
                    emit(i, j, k, mu, wm);
                }
    }
}

// g2p of one particle: fetch(i, j, k, ue[ncomp]) reads the grid; the 12
// components of c_p are returned in cp (row-major 4x3).  A staggered g2p
// leaves zeros in the other components (Picasso_APIC.hpp:505-543).  With
// GRAD (collocated only; the spline must carry gradients) L also receives
// the velocity gradient L(a, b) = sum_n u_n(a) dw_n/dx_b (G2P::gradient,
// Picasso_ParticleInterpolation.hpp:165-181) from the same grid values.
template <int ORDER, bool STAG, bool GRAD = false, typename T, typename Fetch>
__device__ __forceinline__ void apic_g2p_particle(const ApicArgs<T>& a,
                                                  const Spline<ORDER, T>& sd, Fetch&& fetch,
                                                  T cp[12], T* L = nullptr)
{
    constexpr int NK = Knots<ORDER>::NK;
    constexpr int NC = STAG ? 1 : 3;
    static_assert(!(GRAD && STAG), "velocity gradient of collocated fields only");
    // v_p = sum w u and B_p = sum w u d^T (Picasso_APIC.hpp:439-471), summed
    // one dimension at a time: the k sums first (weights w_k and w_k d_k),
    // then j, then i.  The gradient adds the sums with g_k, g_j and g_i.
    T wd[3][NK];
#pragma unroll
    for (int e = 0; e < 3; ++e)
#pragma unroll
        for (int n = 0; n < NK; ++n) wd[e][n] = sd.w[e][n] * sd.d[e][n];
    T up[NC];
    T B[NC][3];
    T G[GRAD ? NC : 1][3];
#pragma unroll
    for (int d = 0; d < NC; ++d) {
        up[d] = T(0);
        B[d][0] = B[d][1] = B[d][2] = T(0);
        if constexpr (GRAD) G[d][0] = G[d][1] = G[d][2] = T(0);
    }
#pragma unroll
    for (int i = 0; i < NK; ++i) {
        T Si[NC], Sdi[NC], Ti[NC], Sgi[NC], Ri[NC];
#pragma unroll
        for (int d = 0; d < NC; ++d) Si[d] = Sdi[d] = Ti[d] = Sgi[d] = Ri[d] = T(0);
#pragma unroll
        for (int j = 0; j < NK; ++j) {
            T sij[NC], tij[NC], rij[NC];
#pragma unroll
            for (int d = 0; d < NC; ++d) sij[d] = tij[d] = rij[d] = T(0);
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                T ue[NC];
                fetch(i, j, k, ue);
#pragma unroll
                for (int d = 0; d < NC; ++d) {
                    sij[d] += sd.w[2][k] * ue[d];
                    tij[d] += wd[2][k] * ue[d];
                    if constexpr (GRAD) rij[d] += sd.g[2][k] * ue[d];
                }
            }
#pragma unroll
            for (int d = 0; d < NC; ++d) {
                Si[d] += sd.w[1][j] * sij[d];
                Sdi[d] += wd[1][j] * sij[d];
                Ti[d] += sd.w[1][j] * tij[d];
                if constexpr (GRAD) {
                    Sgi[d] += sd.g[1][j] * sij[d];
                    Ri[d] += sd.w[1][j] * rij[d];
                }
            }
        }
#pragma unroll
        for (int d = 0; d < NC; ++d) {
            up[d] += sd.w[0][i] * Si[d];
            B[d][0] += wd[0][i] * Si[d];
            B[d][1] += sd.w[0][i] * Sdi[d];
            B[d][2] += sd.w[0][i] * Ti[d];
            if constexpr (GRAD) {
                G[d][0] += sd.g[0][i] * Si[d];
                G[d][1] += sd.w[0][i] * Sgi[d];
                G[d][2] += sd.w[0][i] * Ri[d];
            }
        }
    }
    if constexpr (STAG) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const bool on = d == a.sdim;
            cp[d] = on ? up[0] : T(0);
#pragma unroll
            for (int q = 0; q < 3; ++q) cp[(1 + q) * 3 + d] = on ? B[0][q] : T(0);
        }
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            cp[d] = up[d];
#pragma unroll
            for (int q = 0; q < 3; ++q) cp[(1 + q) * 3 + d] = B[d][q];
        }
        if constexpr (GRAD) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int q = 0; q < 3; ++q) L[d * 3 + q] = G[d][q];
        }
    }
}

template <typename T>
inline ApicArgs<T> make_args(pic_context* ctx, int entity, const pic_mesh& mesh,
                             int64_t np, const pic_pfield* x, const pic_pfield* m,
                             const pic_pfield* c, const pic_grid* gm, const pic_grid& gmu,
                             const uint32_t* perm)
{
    ApicArgs<T> a;
    a.lat = make_lattice(mesh, entity);
    a.np = np;
    a.perm = perm;
    a.list = nullptr;
    a.serial_max = 0u;
    a.x = make_pfield<T>(*x);
    a.m = m ? make_pfield<T>(*m) : PField<T>{};
    a.c = make_pfield<T>(*c);
    a.gm = gm ? make_gview<T>(*gm) : GView<T>{};
    a.gmu = make_gview<T>(gmu);
    a.gnm = GView<T>{};
    a.sdim = stag_dim(entity) < 0 ? 0 : stag_dim(entity);
    a.err = ctx->d_err;
    a.F = a.det = a.Finv = PField<double>{};
    a.dt = 0.0;
    return a;
}

// Order/staggering dispatch to a functor with a templated call operator.
template <typename T, typename Launch>
inline void dispatch_order_stag(int order, bool stag, Launch& L)
{
    switch (order) {
    case 1:
        if (stag) L.template operator()<1, true, T>();
        else L.template operator()<1, false, T>();
        break;
    case 2:
        if (stag) L.template operator()<2, true, T>();
        else L.template operator()<2, false, T>();
        break;
    default:
        if (stag) L.template operator()<3, true, T>();
        else L.template operator()<3, false, T>();
        break;
    }
}

// Longest stale list a deterministic p2g orders (k_sort_stale) and applies
// one particle after another.
constexpr uint32_t STALE_SORT_MAX = 8192;

// Scatter/gather over a device-side slot list (pic_apic.hip).  With
// `serial`, a list of at most STALE_SORT_MAX particles is applied by one
// work-item in list order (longer ones by the whole grid).
int apic_p2g_list(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                  int64_t np, const pic_pfield& x, const pic_pfield& m, const pic_pfield& c,
                  const pic_grid& gm, const pic_grid& gmu, const uint32_t* perm,
                  const uint32_t* list, bool serial = false);
int apic_g2p_list(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                  int64_t np, const pic_pfield& x, const pic_grid& gu, const pic_grid* gmass,
                  const pic_pfield& c, const uint32_t* perm, const uint32_t* list,
                  const FUpdate* fu);

// Brick scheme (pic_apic_brick.hip).  Return PIC_ERR_UNSUPPORTED to fall
// back to the scatter/gather kernels.  `det`: the deterministic p2g (colour
// passes, stale particles one after another in slot order).
int apic_p2g_brick(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                   int64_t np, const pic_pfield& x, const pic_pfield& m, const pic_pfield& c,
                   const pic_grid& gm, const pic_grid& gmu, const pic_bins& bins, bool det);
// Whether the brick kernels take (order, entity) on these bins.
bool apic_brick_supported(int order, int entity, const pic_bins& bins);
int apic_g2p_brick(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                   int64_t np, const pic_pfield& x, const pic_grid& gu, const pic_grid* gmass,
                   const pic_pfield& c, const pic_bins& bins, const FUpdate* fu);
// Sort, zero-first p2g and normalised g2p through the bricks in one pass
// of launches (pic_apic_step); PIC_ERR_UNSUPPORTED when the binning does not
// suit the p2g brick kernels.
int apic_step_bricks(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                     int64_t np, const pic_pfield& x, const pic_pfield& m, const pic_pfield& c,
                     const pic_grid& gm, const pic_grid& gmu, const pic_binning& geom,
                     uint32_t* keys, uint32_t* perm, uint32_t* bin_start, uint32_t* active);

// Attach the F-update outputs to the arguments of a g2p.
template <typename T>
inline void set_F_update(ApicArgs<T>& a, const FUpdate* fu)
{
    if (!fu) return;
    a.F = make_pfield<double>(*fu->F);
    a.det = make_pfield<double>(*fu->det);
    a.Finv = make_pfield<double>(*fu->Finv);
    a.dt = fu->dt;
}

// Grid velocity at a node: the stored value, or momentum / mass (0 where the
// mass is 0) with the division in FP64 when normalisation is fused in.
template <typename T>
__device__ __forceinline__ T node_velocity(const GView<T>& gu, const GView<T>& gm, int64_t off,
                                           int64_t moff, int d)
{
    const T q = gu.data[off + d * gu.cs];
    if (!gm.data) return q;
    const T mm = gm.data[moff];
    return mm > T(0) ? T(double(q) / double(mm)) : T(0);
}

// All three components at a node, with one FP64 reciprocal of the mass.
template <typename T>
__device__ __forceinline__ void node_velocity3(const GView<T>& gu, const GView<T>& gm,
                                               int64_t off, int64_t moff, T u[3])
{
    if (!gm.data) {
#pragma unroll
        for (int d = 0; d < 3; ++d) u[d] = gu.data[off + d * gu.cs];
        return;
    }
    const T mm = gm.data[moff];
    const double r = mm > T(0) ? 1.0 / double(mm) : 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = T(double(gu.data[off + d * gu.cs]) * r);
}

}  // namespace pic
