// Grid halo on one GPU: the gather / scatter-sum halo steps GridOperator
// wraps around every apply (Picasso_GridOperator.hpp:203-215, 465-491),
// for the part of them one GPU owns: periodic dimensions.  Node i of a
// periodic dimension is the image of own_lo + ((i - own_lo) mod own_n).
//  - scatter-sum: every node outside the owned range adds its value into
//    its image, one periodic dimension after the other (an x-ghost's share
//    reaches the owner through the x fold; a corner ghost's through the x,
//    then y, then z folds).  Inside one fold the ghosts and their images
//    are disjoint and no two ghosts share an image (the owned range is at
//    least as wide as both ghost layers together), so each fold is a plain
//    read-modify-write with no atomics and a fixed order;
//  - gather: every node outside the owned range takes its image's value in
//    one pass (the images are owned in every periodic dimension).
// Non-periodic dimensions exchange nothing: on one GPU there is no
// neighbour.  After the scatter, ghost values are unspecified (a fold adds
// into nodes that are ghosts of a later fold); a gather restores them.
#include "pic_common.hpp"
#include "pic_internal.hpp"

namespace pic {

struct HaloGeom {
    int32_t own_lo[3], own_n[3];
    int32_t periodic[3];
};

__device__ __forceinline__ int32_t wrap(int32_t i, int32_t lo, int32_t n)
{
    int32_t r = (i - lo) % n;
    return lo + (r < 0 ? r + n : r);
}

// One fold along dimension D: nodes of the window outside the owned range
// in D add into their image along D.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_halo_fold(GView<T> g, int ncomp, HaloGeom h)
{
    const int64_t nn = int64_t(g.n[0]) * g.n[1] * g.n[2];
    for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < nn;
         q += int64_t(gridDim.x) * blockDim.x) {
        int32_t idx[3] = {g.lo[0] + int32_t(q / (int64_t(g.n[1]) * g.n[2])),
                          g.lo[1] + int32_t((q / g.n[2]) % g.n[1]), g.lo[2] + int32_t(q % g.n[2])};
        const int32_t i = idx[D];
        if (i >= h.own_lo[D] && i < h.own_lo[D] + h.own_n[D]) continue;
        const int64_t src = g.off(idx[0], idx[1], idx[2]);
        idx[D] = wrap(i, h.own_lo[D], h.own_n[D]);
        const int64_t dst = g.off(idx[0], idx[1], idx[2]);
        for (int c = 0; c < ncomp; ++c) g.data[dst + c * g.cs] += g.data[src + c * g.cs];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_halo_gather(GView<T> g, int ncomp, HaloGeom h)
{
    const int64_t nn = int64_t(g.n[0]) * g.n[1] * g.n[2];
    for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < nn;
         q += int64_t(gridDim.x) * blockDim.x) {
        const int32_t idx[3] = {g.lo[0] + int32_t(q / (int64_t(g.n[1]) * g.n[2])),
                                g.lo[1] + int32_t((q / g.n[2]) % g.n[1]),
                                g.lo[2] + int32_t(q % g.n[2])};
        int32_t img[3];
        bool ghost = false;
        for (int d = 0; d < 3; ++d) {
            img[d] = h.periodic[d] ? wrap(idx[d], h.own_lo[d], h.own_n[d]) : idx[d];
            ghost = ghost || img[d] != idx[d];
        }
        if (!ghost) continue;
        const int64_t src = g.off(img[0], img[1], img[2]);
        const int64_t dst = g.off(idx[0], idx[1], idx[2]);
        for (int c = 0; c < ncomp; ++c) g.data[dst + c * g.cs] = g.data[src + c * g.cs];
    }
}

}  // namespace pic

using namespace pic;

extern "C" int pic_grid_halo(pic_context* ctx, int op, int dtype, const pic_grid* g, int ncomp,
                             const int32_t own_lo[3], const int32_t own_n[3],
                             const int32_t periodic[3])
{
    Region region("picasso_amd::grid_halo");
    if (!ctx || !valid_grid(g) || !own_lo || !own_n || !periodic || ncomp < 1 ||
        (op != PIC_HALO_GATHER && op != PIC_HALO_SCATTER_SUM) ||
        (dtype != PIC_F64 && dtype != PIC_F32))
        return PIC_ERR_ARG;
    HaloGeom h;
    for (int d = 0; d < 3; ++d) {
        h.own_lo[d] = own_lo[d];
        h.own_n[d] = own_n[d];
        h.periodic[d] = periodic[d] ? 1 : 0;
        const int64_t wlo = g->lo[d], whi = int64_t(g->lo[d]) + g->n[d];
        const int64_t olo = own_lo[d], ohi = int64_t(own_lo[d]) + own_n[d];
        if (own_n[d] < 1 || olo < wlo || ohi > whi) return PIC_ERR_ARG;
        // One wrap per ghost, no two ghosts on one image.
        if (h.periodic[d] && (olo - wlo) + (whi - ohi) > own_n[d]) return PIC_ERR_ARG;
    }
    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;
    const int64_t nn = int64_t(g->n[0]) * g->n[1] * g->n[2];
    const unsigned b = (unsigned)std::min<int64_t>(blocks_for(nn, 256), 16384);
    TimedLaunch tl(ctx, "grid_halo");
    auto go = [&](auto tag) {
        using T = decltype(tag);
        const GView<T> v = make_gview<T>(*g);
        if (op == PIC_HALO_GATHER) {
            k_halo_gather<T><<<b, 256, 0, ctx->stream>>>(v, ncomp, h);
            return;
        }
        if (h.periodic[0]) k_halo_fold<T, 0><<<b, 256, 0, ctx->stream>>>(v, ncomp, h);
        if (h.periodic[1]) k_halo_fold<T, 1><<<b, 256, 0, ctx->stream>>>(v, ncomp, h);
        if (h.periodic[2]) k_halo_fold<T, 2><<<b, 256, 0, ctx->stream>>>(v, ncomp, h);
    };
    if (dtype == PIC_F64) go(double{});
    else go(float{});
    return launch_status();
}

// ---------------------------------------------------------------------------
// GridOperator's apply for a particle operation on one GPU
// (Picasso_GridOperator.hpp:459-496): gather halos, zeroed scatter fields,
// the operation, scatter-sum halos.

static bool any_periodic(const pic_field_dep& f)
{
    return f.periodic[0] || f.periodic[1] || f.periodic[2];
}

extern "C" int pic_grid_operator_apply(pic_context* ctx, int dtype, int ngather,
                                       const pic_field_dep* gather, int nscatter,
                                       const pic_field_dep* scatter, pic_apply_fn body,
                                       void* user)
{
    Region region("Picasso::GridOperator::apply");
    if (!ctx || !body || ngather < 0 || nscatter < 0 || (ngather > 0 && !gather) ||
        (nscatter > 0 && !scatter) || (dtype != PIC_F64 && dtype != PIC_F32))
        return PIC_ERR_ARG;
    for (int k = 0; k < ngather + nscatter; ++k) {
        const pic_field_dep& f = k < ngather ? gather[k] : scatter[k - ngather];
        if (!valid_grid(&f.grid) || f.ncomp < 1) return PIC_ERR_ARG;
    }
    // Gather distributed dependencies (:465-466): ghosts take their images.
    for (int k = 0; k < ngather; ++k) {
        const pic_field_dep& f = gather[k];
        if (!any_periodic(f)) continue;
        if (const int rc = pic_grid_halo(ctx, PIC_HALO_GATHER, dtype, &f.grid, f.ncomp, f.own_lo,
                                         f.own_n, f.periodic))
            return rc;
    }
    // Scatter fields start from zero (createDependencies, :532-533).
    {
        DeviceGuard dg(ctx->device);
        if (!dg.ok) return PIC_ERR_HIP;
        for (int k = 0; k < nscatter; ++k)
            if (const int rc = zero_grid(ctx, dtype, scatter[k].grid, scatter[k].ncomp)) return rc;
    }
    // The operation (applyOp, :486-487): the caller's transfers.
    if (const int rc = body(ctx, user)) return rc;
    // Scatter distributed dependencies (:489-491): ghosts add into their
    // images.
    for (int k = 0; k < nscatter; ++k) {
        const pic_field_dep& f = scatter[k];
        if (!any_periodic(f)) continue;
        if (const int rc = pic_grid_halo(ctx, PIC_HALO_SCATTER_SUM, dtype, &f.grid, f.ncomp,
                                         f.own_lo, f.own_n, f.periodic))
            return rc;
    }
    return PIC_OK;
}
