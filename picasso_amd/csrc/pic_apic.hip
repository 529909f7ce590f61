// APIC particle<->grid transfers (Picasso_APIC.hpp) for gfx950.
//
// Two execution schemes share the per-particle arithmetic:
//  * scatter: one work-item per particle, hardware float atomics into the
//    grid (any particle order);
//  * brick: particles binned by pic_bin_sort, one workgroup per brick of
//    cells, the brick's node footprint accumulated in LDS and flushed with
//    one global atomic per touched footprint node (pic_apic_brick.hip).
#include "pic_apic.hpp"
#include "pic_internal.hpp"

namespace pic {

// ---------------------------------------------------------------------------
// p2g, scatter scheme.  Grid-stride over particles, or over a device-side
// list of sorted slots (list[0] = count) left by the brick kernels.
template <int ORDER, bool STAG, typename T>
__global__ __launch_bounds__(256) void k_apic_p2g_scatter(ApicArgs<T> a)
{
    const int64_t n = a.list ? int64_t(a.list[0]) : a.np;
    // A short list in deterministic mode: work-item 0 alone, in list order.
    const bool serial = a.list && n <= int64_t(a.serial_max);
    if (serial && (blockIdx.x | threadIdx.x) != 0) return;
    const int64_t stride = serial ? 1 : int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = serial ? 0 : int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += stride) {
        const int64_t s = a.list ? int64_t(a.list[1 + i]) : i;
        const int64_t p = a.perm ? int64_t(a.perm[s]) : s;
        ApicParticle<T> pt;
        load_apic_particle<STAG>(a, p, pt);
        Spline<ORDER, T> sd;
        eval_spline<ORDER, T, ORDER == 1, ORDER != 1>(a.lat, pt.x, sd);
        if (!stencil_inside(sd, a.gmu)) {
            atomicOr(a.err, ERRBIT_BOUNDS);
            continue;
        }
        apic_p2g_particle<ORDER, STAG>(a, pt, sd, [&](int i, int j, int k, const T* mu, T m) {
            i += sd.s0[0];
            j += sd.s0[1];
            k += sd.s0[2];
            const int64_t off = a.gmu.off(i, j, k);
            if constexpr (STAG) {
                atomic_add(a.gmu.data + off, mu[0]);
            } else {
#pragma unroll
                for (int d = 0; d < 3; ++d) atomic_add(a.gmu.data + off + d * a.gmu.cs, mu[d]);
            }
            atomic_add(a.gm.data + a.gm.off(i, j, k), m);
        });
    }
}

// ---------------------------------------------------------------------------
// g2p, gather (any particle order), same iteration scheme.  With GF the
// velocity gradient is formed alongside and applied to F.
template <int ORDER, bool STAG, typename T, bool GF = false>
__global__ __launch_bounds__(256) void k_apic_g2p_gather(ApicArgs<T> a)
{
    const int64_t n = a.list ? int64_t(a.list[0]) : a.np;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t s = a.list ? int64_t(a.list[1 + i]) : i;
        const int64_t p = a.perm ? int64_t(a.perm[s]) : s;
        const int64_t pb = a.x.base(p);
        T xp[3] = {a.x.at(pb, 0), a.x.at(pb, 1), a.x.at(pb, 2)};
        Spline<ORDER, T> sd;
        eval_spline<ORDER, T, GF, true>(a.lat, xp, sd);
        if (!stencil_inside(sd, a.gmu)) {
            atomicOr(a.err, ERRBIT_BOUNDS);
            continue;
        }
        T cp[12], L[9];
        apic_g2p_particle<ORDER, STAG, GF>(a, sd, [&](int i, int j, int k, T* ue) {
            const int ni = sd.s0[0] + i, nj = sd.s0[1] + j, nk = sd.s0[2] + k;
            const int64_t off = a.gmu.off(ni, nj, nk);
            const int64_t moff = a.gnm.data ? a.gnm.off(ni, nj, nk) : 0;
            if constexpr (STAG) {
                ue[0] = node_velocity(a.gmu, a.gnm, off, moff, 0);
            } else {
#pragma unroll
                for (int d = 0; d < 3; ++d) ue[d] = node_velocity(a.gmu, a.gnm, off, moff, d);
            }
        }, cp, L);
        const int64_t cb = a.c.base(p);
#pragma unroll
        for (int r = 0; r < 12; ++r) a.c.at(cb, r) = cp[r];
        if constexpr (GF) store_F_update(a, p, L);
    }
}

// ---------------------------------------------------------------------------
// Host dispatch.  With a slot list the element count lives on the device,
// so a fixed grid strides over it.
constexpr unsigned LIST_BLOCKS = 512;

static unsigned launch_blocks(const void* list, int64_t np)
{
    return list ? LIST_BLOCKS : blocks_for(np, 256);
}

template <typename T>
struct P2GScatterLaunch {
    ApicArgs<T> a;
    hipStream_t st;
    template <int O, bool S, typename U>
    void operator()()
    {
        k_apic_p2g_scatter<O, S, U><<<launch_blocks(a.list, a.np), 256, 0, st>>>(a);
    }
};

template <typename T>
struct G2PGatherLaunch {
    ApicArgs<T> a;
    hipStream_t st;
    template <int O, bool S, typename U>
    void operator()()
    {
        if constexpr (!S) {
            if (a.F.data) {
                k_apic_g2p_gather<O, S, U, true><<<launch_blocks(a.list, a.np), 256, 0, st>>>(a);
                return;
            }
        }
        k_apic_g2p_gather<O, S, U><<<launch_blocks(a.list, a.np), 256, 0, st>>>(a);
    }
};

int apic_p2g_list(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                  int64_t np, const pic_pfield& x, const pic_pfield& m, const pic_pfield& c,
                  const pic_grid& gm, const pic_grid& gmu, const uint32_t* perm,
                  const uint32_t* list, bool serial)
{
    const bool stag = stag_dim(entity) >= 0;
    TimedLaunch tl(ctx, "apic_p2g_stale");
    if (dtype == PIC_F64) {
        P2GScatterLaunch<double> L{make_args<double>(ctx, entity, mesh, np, &x, &m, &c, &gm, gmu,
                                                     perm),
                                   ctx->stream};
        L.a.list = list;
        L.a.serial_max = serial ? STALE_SORT_MAX : 0u;
        dispatch_order_stag<double>(order, stag, L);
    } else {
        P2GScatterLaunch<float> L{make_args<float>(ctx, entity, mesh, np, &x, &m, &c, &gm, gmu,
                                                   perm),
                                  ctx->stream};
        L.a.list = list;
        L.a.serial_max = serial ? STALE_SORT_MAX : 0u;
        dispatch_order_stag<float>(order, stag, L);
    }
    return launch_status();
}

int apic_g2p_list(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                  int64_t np, const pic_pfield& x, const pic_grid& gu, const pic_grid* gmass,
                  const pic_pfield& c, const uint32_t* perm, const uint32_t* list,
                  const FUpdate* fu)
{
    const bool stag = stag_dim(entity) >= 0;
    TimedLaunch tl(ctx, "apic_g2p_stale");
    if (dtype == PIC_F64) {
        G2PGatherLaunch<double> L{
            make_args<double>(ctx, entity, mesh, np, &x, nullptr, &c, nullptr, gu, perm),
            ctx->stream};
        L.a.list = list;
        if (gmass) L.a.gnm = make_gview<double>(*gmass);
        set_F_update(L.a, fu);
        dispatch_order_stag<double>(order, stag, L);
    } else {
        G2PGatherLaunch<float> L{
            make_args<float>(ctx, entity, mesh, np, &x, nullptr, &c, nullptr, gu, perm),
            ctx->stream};
        L.a.list = list;
        if (gmass) L.a.gnm = make_gview<float>(*gmass);
        set_F_update(L.a, fu);
        dispatch_order_stag<float>(order, stag, L);
    }
    return launch_status();
}

}  // namespace pic

using namespace pic;

static bool same_window(const pic_grid& a, const pic_grid& b)
{
    for (int d = 0; d < 3; ++d)
        if (a.lo[d] != b.lo[d] || a.n[d] != b.n[d]) return false;
    return true;
}

extern "C" int pic_apic_p2g(pic_context* ctx, int order, int entity, int dtype,
                            const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                            const pic_pfield* m_p, const pic_pfield* c_p,
                            const pic_grid* m_i, const pic_grid* mu_i,
                            const pic_bins* bins, uint32_t flags)
{
    Region region("picasso_amd::APIC::p2g");
    if (!ctx || !mesh || !(mesh->h > 0.0) || order < 1 || order > 3 ||
        !valid_entity(entity) || (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 ||
        np > int64_t(UINT32_MAX) || !valid_grid(m_i) || !valid_grid(mu_i) ||
        !same_window(*m_i, *mu_i))
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !valid_pfield(m_p) || !valid_pfield(c_p)))
        return PIC_ERR_ARG;
    if (bins && !bins->perm) return PIC_ERR_ARG;
    const bool det = (flags & PIC_P2G_DETERMINISTIC) != 0;
    if (det && !(bins && apic_brick_supported(order, entity, *bins))) return PIC_ERR_UNSUPPORTED;
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    const bool stag = stag_dim(entity) >= 0;
    if (flags & PIC_P2G_ZERO_FIRST) {
        int rc = zero_grid(ctx, dtype, *m_i, 1);
        if (rc) return rc;
        rc = zero_grid(ctx, dtype, *mu_i, stag ? 1 : 3);
        if (rc) return rc;
    }
    if (np == 0) return PIC_OK;
    if (bins) {
        int rc = apic_p2g_brick(ctx, order, entity, dtype, *mesh, np, *x, *m_p, *c_p,
                                *m_i, *mu_i, *bins, det);
        if (rc != PIC_ERR_UNSUPPORTED) return rc;
    }
    // The per-particle kernel's atomics land in any order.
    if (det) return PIC_ERR_UNSUPPORTED;
    TimedLaunch tl(ctx, "apic_p2g");
    const uint32_t* perm = bins ? bins->perm : nullptr;
    if (dtype == PIC_F64) {
        P2GScatterLaunch<double> L{make_args<double>(ctx, entity, *mesh, np, x, m_p, c_p, m_i,
                                                     *mu_i, perm),
                                   ctx->stream};
        dispatch_order_stag<double>(order, stag, L);
    } else {
        P2GScatterLaunch<float> L{make_args<float>(ctx, entity, *mesh, np, x, m_p, c_p, m_i,
                                                   *mu_i, perm),
                                  ctx->stream};
        dispatch_order_stag<float>(order, stag, L);
    }
    return launch_status();
}

static int apic_g2p_impl(pic_context* ctx, int order, int entity, int dtype,
                         const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                         const pic_grid* u_i, const pic_grid* m_i, const pic_pfield* c_p,
                         const pic_bins* bins, const FUpdate* fu = nullptr)
{
    if (!ctx || !mesh || !(mesh->h > 0.0) || order < 1 || order > 3 ||
        !valid_entity(entity) || (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 ||
        np > int64_t(UINT32_MAX) || !valid_grid(u_i) || (m_i && !valid_grid(m_i)) ||
        (m_i && !same_window(*m_i, *u_i)))
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !valid_pfield(c_p))) return PIC_ERR_ARG;
    if (fu && ((entity != PIC_NODE && entity != PIC_CELL) ||
               (np > 0 && (!valid_pfield(fu->F) || !valid_pfield(fu->det) ||
                           !valid_pfield(fu->Finv)))))
        return PIC_ERR_ARG;
    if (bins && !bins->perm) return PIC_ERR_ARG;
    if (np == 0) return PIC_OK;
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    const bool stag = stag_dim(entity) >= 0;
    if (bins) {
        int rc = apic_g2p_brick(ctx, order, entity, dtype, *mesh, np, *x, *u_i, m_i, *c_p, *bins,
                                fu);
        if (rc != PIC_ERR_UNSUPPORTED) return rc;
    }
    TimedLaunch tl(ctx, "apic_g2p");
    const uint32_t* perm = bins ? bins->perm : nullptr;
    if (dtype == PIC_F64) {
        G2PGatherLaunch<double> L{
            make_args<double>(ctx, entity, *mesh, np, x, nullptr, c_p, nullptr, *u_i, perm),
            ctx->stream};
        if (m_i) L.a.gnm = make_gview<double>(*m_i);
        set_F_update(L.a, fu);
        dispatch_order_stag<double>(order, stag, L);
    } else {
        G2PGatherLaunch<float> L{
            make_args<float>(ctx, entity, *mesh, np, x, nullptr, c_p, nullptr, *u_i, perm),
            ctx->stream};
        if (m_i) L.a.gnm = make_gview<float>(*m_i);
        set_F_update(L.a, fu);
        dispatch_order_stag<float>(order, stag, L);
    }
    return launch_status();
}

extern "C" int pic_apic_g2p(pic_context* ctx, int order, int entity, int dtype,
                            const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                            const pic_grid* u_i, const pic_pfield* c_p,
                            const pic_bins* bins)
{
    Region region("picasso_amd::APIC::g2p");
    return apic_g2p_impl(ctx, order, entity, dtype, mesh, np, x, u_i, nullptr, c_p, bins);
}

extern "C" int pic_apic_g2p_normalized(pic_context* ctx, int order, int entity, int dtype,
                                       const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                                       const pic_grid* m_i, const pic_grid* mu_i,
                                       const pic_pfield* c_p, const pic_bins* bins)
{
    Region region("picasso_amd::APIC::g2p_normalized");
    if (!m_i) return PIC_ERR_ARG;
    return apic_g2p_impl(ctx, order, entity, dtype, mesh, np, x, mu_i, m_i, c_p, bins);
}

extern "C" int pic_apic_step(pic_context* ctx, int order, int entity, int dtype,
                             const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                             const pic_pfield* m_p, const pic_pfield* c_p, const pic_grid* m_i,
                             const pic_grid* mu_i, const pic_binning* geom, uint32_t* keys,
                             uint32_t* perm, uint32_t* bin_start, uint32_t* active)
{
    Region region("picasso_amd::APIC::step");
    if (!ctx || !mesh || !(mesh->h > 0.0) || order < 1 || order > 3 ||
        !valid_entity(entity) || (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 ||
        np >= int64_t(UINT32_MAX) || !valid_grid(m_i) || !valid_grid(mu_i) ||
        !same_window(*m_i, *mu_i) || !geom || !bin_start)
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !valid_pfield(m_p) || !valid_pfield(c_p) || !keys ||
                   !perm))
        return PIC_ERR_ARG;
    const int64_t nb = pic_bin_count(geom);
    if (nb < 0) return int(nb);
    {
        DeviceGuard g(ctx->device);
        if (!g.ok) return PIC_ERR_HIP;
        const int rc = apic_step_bricks(ctx, order, entity, dtype, *mesh, np, *x, *m_p, *c_p,
                                        *m_i, *mu_i, *geom, keys, perm, bin_start, active);
        if (rc != PIC_ERR_UNSUPPORTED) return rc;
    }
    // Binnings the p2g bricks do not take, and empty particle sets: the
    // three calls one after another.
    if (const int rc = pic_bin_sort(ctx, dtype, mesh, np, x, geom, keys, perm, bin_start, active))
        return rc;
    const pic_bins bins{perm, keys, bin_start, active, *geom};
    const pic_bins* bp = np > 0 ? &bins : nullptr;  // no particles: no perm either
    if (const int rc = pic_apic_p2g(ctx, order, entity, dtype, mesh, np, x, m_p, c_p, m_i, mu_i,
                                    bp, PIC_P2G_ZERO_FIRST))
        return rc;
    return pic_apic_g2p_normalized(ctx, order, entity, dtype, mesh, np, x, m_i, mu_i, c_p, bp);
}

extern "C" int pic_apic_g2p_update_F(pic_context* ctx, int order, int entity, int dtype,
                                     const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                                     const pic_grid* m_i, const pic_grid* u_i,
                                     const pic_pfield* c_p, double dt, const pic_pfield* F,
                                     const pic_pfield* det, const pic_pfield* Finv,
                                     const pic_bins* bins)
{
    Region region("picasso_amd::APIC::g2p_update_F");
    const FUpdate fu{F, det, Finv, dt};
    return apic_g2p_impl(ctx, order, entity, dtype, mesh, np, x, u_i, m_i, c_p, bins, &fu);
}
