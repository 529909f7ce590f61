// APIC brick kernels (single instance) and their host launchers; the
// per-brick device code is in pic_apic_brick.hpp.
#include "pic_apic_brick.hpp"

namespace pic {

template <int ORDER, typename T, bool STAG>
__global__ __launch_bounds__((P2GShape<ORDER, T>::NT), (P2GShape<ORDER, T>::WAVES)) void k_apic_p2g_brick(
    BrickArgs<T> g)
{
    using S = P2GShape<ORDER, T>;
    __shared__ __align__(16) T lds[S::LDS];
    __shared__ uint32_t rng[2 * 2 * S::BINS];
    __shared__ BrickItem q[2];
    for (int i = threadIdx.x; i < 4 * S::BINS; i += S::NT) rng[i] = 0u;
    P2GState<ORDER, T> st{rng, 0, false, {}};
    walk_bricks<true>(g, q, [&](const BrickItem& it, BrickWalk<true, BrickArgs<T>>& w) {
        p2g_one_brick<ORDER, T, STAG>(g, it, w, lds, st);
    });
}

// Deterministic p2g, one launch per brick colour: the bricks of colour c
// (grouped by k_colour_items into `items`, counts and offsets in `colour`)
// under a static walk, flushed with plain read-modify-writes.  No two
// bricks of a colour share a node, so every node receives its partial sums
// in colour order whatever the scheduling.
template <int ORDER, typename T, bool STAG>
__global__ __launch_bounds__((P2GShape<ORDER, T>::NT), (P2GShape<ORDER, T>::WAVES)) void
k_apic_p2g_brick_colour(BrickArgs<T> g, const BrickItem* items, const uint32_t* colour, uint32_t c)
{
    using S = P2GShape<ORDER, T>;
    __shared__ __align__(16) T lds[S::LDS];
    __shared__ uint32_t rng[2 * 2 * S::BINS];
    __shared__ BrickItem q[2];
    for (int i = threadIdx.x; i < 4 * S::BINS; i += S::NT) rng[i] = 0u;
    BrickArgs<T> gc = g;
    gc.items = items + colour[8 + c];
    gc.active = colour + c;  // active[0]: this colour's brick count
    P2GState<ORDER, T> st{rng, 0, false, {}};
    walk_bricks<false>(gc, q, [&](const BrickItem& it, BrickWalk<false, BrickArgs<T>>& w) {
        p2g_one_brick<ORDER, T, STAG, true, false>(gc, it, w, lds, st);
    });
}

// Stale list [count, slots...] sorted ascending in place (one workgroup,
// bitonic in LDS), so that applying its particles one after another gives
// the same sums on every run.  Longer lists than STALE_SORT_MAX stay as
// they are and raise ERRBIT_CAPACITY.
__global__ __launch_bounds__(1024) void k_sort_stale(uint32_t* list, int* err)
{
    __shared__ uint32_t s[STALE_SORT_MAX];
    const uint32_t n = list[0];
    if (n <= 1u) return;
    if (n > STALE_SORT_MAX) {
        if (threadIdx.x == 0) atomicOr(err, ERRBIT_CAPACITY);
        return;
    }
    uint32_t m = 2;
    while (m < n) m <<= 1;
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) s[i] = i < n ? list[1 + i] : ~0u;
    __syncthreads();
    for (uint32_t k = 2; k <= m; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
                const uint32_t ij = i ^ j;
                if (ij > i) {
                    const uint32_t a = s[i], b = s[ij];
                    if ((a > b) == ((i & k) == 0)) {
                        s[i] = b;
                        s[ij] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) list[1 + i] = s[i];
}

template <int BB, int ORDER, typename T, bool GF, bool STAG, int NT, int R>
__global__ __launch_bounds__(NT) void k_apic_g2p_brick(BrickArgs<T> g)
{
    constexpr int FN = BB + ORDER;
    __shared__ T us[(STAG ? 1 : 3) * FN * FN * FN];
    __shared__ BrickItem q[2];
    walk_bricks<true>(g, q, [&](const BrickItem& it, BrickWalk<true, BrickArgs<T>>& w) {
        g2p_one_brick<BB, ORDER, T, GF, STAG, NT, R, true>(g, it, w, us);
    });
}

template <typename T>
static int make_brick_args(pic_context* ctx, int order, int entity, const pic_mesh& mesh,
                            int64_t np, const pic_pfield& x, const pic_pfield* m,
                            const pic_pfield& c, const pic_grid* gm, const pic_grid& gmu,
                            const pic_grid* gmass, const pic_bins& bins, BrickArgs<T>& g,
                            size_t extra_bytes = 0, void** extra = nullptr)
{
    const int rc = setup_walk(ctx, order, entity, np, bins, g, extra_bytes, extra);
    if (rc) return rc;
    g.a = make_args<T>(ctx, entity, mesh, np, &x, m, &c, gm, gmu, bins.perm);
    if (gmass) g.a.gnm = make_gview<T>(*gmass);
    return PIC_OK;
}

template <typename T>
struct P2GBrickLaunch {
    pic_context* ctx;
    const BrickArgs<T>& g;
    bool stag;
    // Deterministic: the colour passes over `items` / `colour`.
    const BrickItem* items = nullptr;
    const uint32_t* colour = nullptr;
    template <int O, bool STAG>
    void launch()
    {
        constexpr int NT = P2GShape<O, T>::NT;
        if (items) {
            const unsigned nbk =
                persistent_blocks(ctx, k_apic_p2g_brick_colour<O, T, STAG>, NT, 0, g.nbricks);
            for (uint32_t c = 0; c < 8; ++c)
                k_apic_p2g_brick_colour<O, T, STAG><<<nbk, NT, 0, ctx->stream>>>(g, items, colour, c);
            return;
        }
        const unsigned nbk =
            persistent_blocks(ctx, k_apic_p2g_brick<O, T, STAG>, NT, 0, g.nbricks);
        k_apic_p2g_brick<O, T, STAG><<<nbk, NT, 0, ctx->stream>>>(g);
    }
    template <int O>
    void go()
    {
        if (stag) launch<O, true>();
        else launch<O, false>();
    }
    void by_order(int order)
    {
        if (order == 1) go<1>();
        else if (order == 2) go<2>();
        else go<3>();
    }
};

template <typename T>
struct G2PBrickLaunch {
    pic_context* ctx;
    const BrickArgs<T>& g;
    bool stag;
    template <int BB, int O, bool GF, bool STAG, int NT, int R>
    void launch_shape()
    {
        const unsigned nbk =
            persistent_blocks(ctx, k_apic_g2p_brick<BB, O, T, GF, STAG, NT, R>, NT, 0, g.nbricks);
        k_apic_g2p_brick<BB, O, T, GF, STAG, NT, R><<<nbk, NT, 0, ctx->stream>>>(g);
    }
    // The collocated kernel without the F update runs either workgroup
    // shape (pic_set_option); the others keep their measured one.
    template <int BB, int O, bool GF, bool STAG>
    void launch()
    {
        using S = G2PShape<T, GF, STAG>;
        if constexpr (!GF && !STAG) {
            if (ctx->g2p_group == 128) return launch_shape<BB, O, GF, STAG, 128, 4>();
            if (ctx->g2p_group == 256) return launch_shape<BB, O, GF, STAG, BT, 3>();
        }
        launch_shape<BB, O, GF, STAG, S::NT, S::R>();
    }
    // The F update is collocated only (the caller rejects it otherwise).
    template <int BB, int O>
    void go()
    {
        if (stag) launch<BB, O, false, true>();
        else if (g.a.F.data) launch<BB, O, true, false>();
        else launch<BB, O, false, false>();
    }
    template <int BB>
    void by_order(int order)
    {
        if (order == 1) go<BB, 1>();
        else if (order == 2) go<BB, 2>();
        else go<BB, 3>();
    }
};

template <typename T>
static int p2g_brick_run(pic_context* ctx, int order, int entity, const pic_mesh& mesh, int64_t np,
                         const pic_pfield& x, const pic_pfield& m, const pic_pfield& c,
                         const pic_grid& gm, const pic_grid& gmu, const pic_bins& bins,
                         int dtype, bool det)
{
    if (brick_edge(bins.geom) != PB) return PIC_ERR_UNSUPPORTED;
    BrickArgs<T> g;
    // Deterministic: 16 words of colour counts and offsets, then the walk
    // items grouped by colour.
    const int64_t nb = pic_bin_count(&bins.geom);
    void* extra = nullptr;
    const size_t extra_bytes = det && nb > 0 ? 64 + size_t(nb) * sizeof(BrickItem) : 0;
    if (const int rc = make_brick_args<T>(ctx, order, entity, mesh, np, x, &m, c, &gm, gmu, nullptr,
                                          bins, g, extra_bytes, &extra))
        return rc;
    P2GBrickLaunch<T> L{ctx, g, stag_dim(entity) >= 0};
    if (det) {
        L.colour = static_cast<uint32_t*>(extra);
        L.items = reinterpret_cast<const BrickItem*>(static_cast<char*>(extra) + 64);
        k_colour_items<<<1, 1024, 0, ctx->stream>>>(g, const_cast<BrickItem*>(L.items),
                                                     const_cast<uint32_t*>(L.colour));
        if (launch_status()) return PIC_ERR_HIP;
    }
    {
        TimedLaunch tl(ctx, "apic_p2g");
        L.by_order(order);
        if (launch_status()) return PIC_ERR_HIP;
    }
    if (det) {
        // Stale particles one after another in slot order.
        k_sort_stale<<<1, 1024, 0, ctx->stream>>>(g.stale, ctx->d_err);
        if (launch_status()) return PIC_ERR_HIP;
    }
    return apic_p2g_list(ctx, order, entity, dtype, mesh, np, x, m, c, gm, gmu, bins.perm, g.stale,
                         det);
}

template <typename T>
static int g2p_brick_run(pic_context* ctx, int order, int entity, const pic_mesh& mesh, int64_t np,
                         const pic_pfield& x, const pic_grid& gu, const pic_grid* gmass,
                         const pic_pfield& c, const pic_bins& bins, int dtype, const FUpdate* fu)
{
    if (fu && stag_dim(entity) >= 0) return PIC_ERR_UNSUPPORTED;
    BrickArgs<T> g;
    if (const int rc = make_brick_args<T>(ctx, order, entity, mesh, np, x, nullptr, c, nullptr, gu,
                                          gmass, bins, g))
        return rc;
    set_F_update(g.a, fu);
    {
        TimedLaunch tl(ctx, "apic_g2p");
        G2PBrickLaunch<T> L{ctx, g, stag_dim(entity) >= 0};
        if (brick_edge(bins.geom) == 4) L.template by_order<4>(order);
        else L.template by_order<8>(order);
        if (launch_status()) return PIC_ERR_HIP;
    }
    return apic_g2p_list(ctx, order, entity, dtype, mesh, np, x, gu, gmass, c, bins.perm, g.stale,
                         fu);
}

// One APIC step through the bricks (pic_apic_step): the sort with the grid
// windows, the walk counters and both stale counts cleared by its key pass
// and the walk items listed by its scan, then the p2g and normalised g2p
// brick kernels and their stale lists.  Scratch: the sort's layout, then
// p2g walk counters | g2p walk counters | p2g stale list (1 + np words) |
// g2p stale list | items.
template <typename T>
static int step_transfers(pic_context* ctx, int order, int entity, int dtype,
                          const pic_mesh& mesh, int64_t np, const pic_pfield& x,
                          const pic_pfield& m, const pic_pfield& c, const pic_grid& gm,
                          const pic_grid& gmu, const pic_bins& bins, uint32_t* ctr,
                          uint32_t* stale_p2g, uint32_t* stale_g2p, const BrickItem* items)
{
    const bool stag = stag_dim(entity) >= 0;
    BrickArgs<T> g;
    if (const int rc = walk_geometry(order, entity, bins, g)) return rc;
    g.items = items;
    g.walk = ctr;
    g.stale = stale_p2g;
    g.a = make_args<T>(ctx, entity, mesh, np, &x, &m, &c, &gm, gmu, bins.perm);
    {
        TimedLaunch tl(ctx, "apic_p2g");
        P2GBrickLaunch<T> L{ctx, g, stag};
        L.by_order(order);
        if (launch_status()) return PIC_ERR_HIP;
    }
    if (const int rc = apic_p2g_list(ctx, order, entity, dtype, mesh, np, x, m, c, gm, gmu,
                                     bins.perm, stale_p2g, false))
        return rc;
    BrickArgs<T> h = g;
    h.walk = ctr + size_t(WALK_LANES) * WALK_PAD;
    h.stale = stale_g2p;
    h.a = make_args<T>(ctx, entity, mesh, np, &x, nullptr, &c, nullptr, gmu, bins.perm);
    h.a.gnm = make_gview<T>(gm);
    {
        TimedLaunch tl(ctx, "apic_g2p");
        G2PBrickLaunch<T> L{ctx, h, stag};
        L.template by_order<PB>(order);
        if (launch_status()) return PIC_ERR_HIP;
    }
    return apic_g2p_list(ctx, order, entity, dtype, mesh, np, x, gmu, &gm, c, bins.perm,
                         stale_g2p, nullptr);
}

int apic_step_bricks(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                     int64_t np, const pic_pfield& x, const pic_pfield& m, const pic_pfield& c,
                     const pic_grid& gm, const pic_grid& gmu, const pic_binning& geom,
                     uint32_t* keys, uint32_t* perm, uint32_t* bin_start, uint32_t* active)
{
    int32_t knot_off[3];
    const int64_t nb = pic_bin_count(&geom);
    if (np <= 0 || nb <= 0 || brick_edge(geom) != PB ||
        !aligned_knot_offset(order, entity, geom.dual, knot_off))
        return PIC_ERR_UNSUPPORTED;
    const SortLayout L = sort_layout(nb, np, active != nullptr);
    const size_t sort_bytes = (L.bytes + 255) & ~size_t(255);
    const size_t ctr_words = size_t(WALK_LANES) * WALK_PAD;
    const size_t stale_words = 1 + size_t(np);
    // p2g's walk counters, g2p's, then the two stale lists
    const size_t items_at =
        (sort_bytes + (2 * ctr_words + 2 * stale_words) * sizeof(uint32_t) + 15) & ~size_t(15);
    char* base = static_cast<char*>(scratch(ctx, items_at + size_t(nb) * sizeof(BrickItem)));
    if (!base) return PIC_ERR_ALLOC;
    uint32_t* sc = reinterpret_cast<uint32_t*>(base);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(base + sort_bytes);
    uint32_t* stale_p2g = ctr + 2 * ctr_words;
    uint32_t* stale_g2p = stale_p2g + stale_words;
    BrickItem* items = reinterpret_cast<BrickItem*>(base + items_at);

    // What the key pass clears: dense grid windows (others get a fill of
    // their own first), the walk counters and the two stale counts.
    const size_t esz = dtype == PIC_F64 ? 8 : 4;
    const int ncomp = stag_dim(entity) >= 0 ? 1 : 3;
    ZeroJob z;
    const pic_grid* grids[2] = {&gm, &gmu};
    const int gcomp[2] = {1, ncomp};
    for (int k = 0; k < 2; ++k) {
        int64_t count = 0;
        if (!(dense_window(*grids[k], gcomp[k], &count) &&
              z.add(grids[k]->data, count * int64_t(esz)))) {
            if (const int rc = zero_grid(ctx, dtype, *grids[k], gcomp[k])) return rc;
        }
    }
    z.add(ctr, int64_t(2 * ctr_words * sizeof(uint32_t)));
    z.add(stale_p2g, sizeof(uint32_t));
    z.add(stale_g2p, sizeof(uint32_t));

    uint32_t* act = active ? active : sort_active(L, sc);
    {
        TimedLaunch tl(ctx, "bin_sort");
        if (const int rc = bin_sort_with(ctx, dtype, mesh, np, &x, geom, nb, L, sc, keys, perm,
                                         bin_start, active, z, items))
            return rc;
    }
    const pic_bins bins{perm, keys, bin_start, act, geom};
    if (dtype == PIC_F32)
        return step_transfers<float>(ctx, order, entity, dtype, mesh, np, x, m, c, gm, gmu, bins,
                                     ctr, stale_p2g, stale_g2p, items);
    return step_transfers<double>(ctx, order, entity, dtype, mesh, np, x, m, c, gm, gmu, bins,
                                  ctr, stale_p2g, stale_g2p, items);
}

bool apic_brick_supported(int order, int entity, const pic_bins& bins)
{
    int32_t knot_off[3];
    return bins.perm && bins.keys && bins.bin_start && brick_edge(bins.geom) == PB &&
           pic_bin_count(&bins.geom) > 0 && aligned_knot_offset(order, entity, bins.geom.dual, knot_off);
}

int apic_p2g_brick(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                   int64_t np, const pic_pfield& x, const pic_pfield& m, const pic_pfield& c,
                   const pic_grid& gm, const pic_grid& gmu, const pic_bins& bins, bool det)
{
    if (dtype == PIC_F32)
        return p2g_brick_run<float>(ctx, order, entity, mesh, np, x, m, c, gm, gmu, bins, dtype,
                                    det);
    return p2g_brick_run<double>(ctx, order, entity, mesh, np, x, m, c, gm, gmu, bins, dtype, det);
}

int apic_g2p_brick(pic_context* ctx, int order, int entity, int dtype, const pic_mesh& mesh,
                   int64_t np, const pic_pfield& x, const pic_grid& gu, const pic_grid* gmass,
                   const pic_pfield& c, const pic_bins& bins, const FUpdate* fu)
{
    if (dtype == PIC_F32)
        return g2p_brick_run<float>(ctx, order, entity, mesh, np, x, gu, gmass, c, bins, dtype, fu);
    return g2p_brick_run<double>(ctx, order, entity, mesh, np, x, gu, gmass, c, bins, dtype, fu);
}

}  // namespace pic
