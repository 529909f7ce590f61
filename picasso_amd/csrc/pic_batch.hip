// Batches of independent APIC instances (no reference counterpart: Picasso
// runs one simulation per GridOperator and launches one Kokkos kernel per
// operator, Picasso_GridOperator.hpp:632-641).  Every pass of a step is one
// launch over all instances:
//  - bins: instance k's bricks are numbered k * S + b (S the smallest power
//    of two holding the largest instance's bricks), so one stable counting
//    sort over all particles bins every instance at once and a brick's
//    instance is its id shifted right by log2 S;
//  - grid zeroing: one kernel over a table of 64 KiB chunks of every
//    instance's windows;
//  - p2g and g2p: the brick kernels' per-brick bodies (pic_apic_brick.hpp)
//    walk the non-empty bricks of all instances; each brick reads its
//    instance's lattice, bins and grid windows from the instance table;
//  - particles whose stencil is not where their bin says take per-particle
//    kernels that find their instance by binary search.
// The batch owns its bins, walk lists and scratch, so its calls allocate
// nothing and a whole step can be captured in one graph.
#include <vector>

#include "pic_apic_brick.hpp"

// Device record of one instance.
struct BatchInst {
    pic::KeyGeom kg;      // its bins
    pic::Lattice lat;     // its node lattice
    int64_t first, np;    // its particles
    uint32_t brick_base;  // global id of its brick 0 (k * S)
    int32_t knot_off[3];
    pic_grid m_i, mu_i;   // its grid windows
};

// A piece of a grid window to zero: n 4-byte words at p, stored 16 bytes at
// a time when `wide` (p 16-byte aligned, n a multiple of 4).
struct ZeroChunk {
    uint32_t* p;
    uint32_t n;
    uint32_t wide;
};

struct pic_batch {
    int device = 0, order = 2, dtype = PIC_F32, ninst = 0;
    int64_t np = 0;       // particles of all instances
    int64_t nb = 0;       // global bricks (ninst << shift)
    uint32_t shift = 0;
    int64_t nchunks = 0;
    BatchInst* inst = nullptr;
    int64_t* first = nullptr;  // ninst + 1 entries
    ZeroChunk* chunks = nullptr;
    uint32_t *keys = nullptr, *perm = nullptr, *bin_start = nullptr, *active = nullptr;
    pic::BrickItem* items = nullptr;
    uint32_t* walk = nullptr;   // walk counters, then the stale count and slots
    uint32_t* sort_sc = nullptr;
    pic::SortLayout L{};
    std::vector<void*> allocs;
};

namespace pic {

template <typename T>
__device__ __forceinline__ GView<T> grid_view(const pic_grid& g)
{
    GView<T> v;
    v.data = static_cast<T*>(g.data);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        v.lo[d] = g.lo[d];
        v.n[d] = g.n[d];
        v.s[d] = g.stride[d];
    }
    v.cs = g.cstride;
    return v;
}

// Instance of particle p: the largest k with first[k] <= p.
__device__ __forceinline__ int instance_of(const int64_t* first, int ninst, int64_t p)
{
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Key functor of the batch: instance k's local key and brick, moved to its
// global bricks.  begin() finds the instance of the workgroup's first
// particle; each particle then steps from there (instances are contiguous).
// A workgroup whose particles all lie in that instance (all but the few
// that straddle an instance boundary) skips the per-particle steps and
// reads the instance's record with scalar loads.
template <typename T>
struct BatchKeys {
    PField<T> x;
    const BatchInst* inst;
    const int64_t* first;
    int ninst;
    uint32_t bvol;
    int k0;
    bool one;      // particles [p_lo, p_lo + SORT_PT * SORT_ITEMS) all in instance k0
    int64_t p_lo;

    __device__ __forceinline__ void begin(int64_t p0)
    {
        k0 = instance_of(first, ninst, p0);
        p_lo = p0;
        one = k0 + 1 >= ninst || first[k0 + 1] >= p0 + SORT_PT * SORT_ITEMS;
    }

    __device__ __forceinline__ uint32_t key_in(const BatchInst& I, int64_t p, uint32_t* b) const
    {
        uint32_t key = cell_key(I.kg, x, p, b);
        if (key != 0xffffffffu) {
            *b += I.brick_base;
            key += I.brick_base * bvol;
        }
        return key;
    }

    __device__ __forceinline__ uint32_t operator()(int64_t p, uint32_t* b) const
    {
        if (one && p >= p_lo) return key_in(inst[k0], p, b);
        int k = k0;
        while (k > 0 && first[k] > p) --k;
        while (k + 1 < ninst && first[k + 1] <= p) ++k;
        // One instance for the whole wave (all but the waves straddling a
        // boundary): read its record with uniform loads.
        const int ku = __builtin_amdgcn_readfirstlane(k);
        if (__all(k == ku)) return key_in(inst[ku], p, b);
        return key_in(inst[k], p, b);
    }
};

// The instance's lattice, bins and grids over the batch-wide walk fields.
template <typename T>
__device__ __forceinline__ void instance_args(const BatchInst& I, BrickArgs<T>& g, bool normalized)
{
    g.a.lat = I.lat;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        g.cell_lo[d] = I.kg.cell_lo[d];
        g.nb[d] = I.kg.nb[d];
        g.knot_off[d] = I.knot_off[d];
    }
    g.brick_base = I.brick_base;
    if (normalized) {
        g.a.gmu = grid_view<T>(I.mu_i);
        g.a.gnm = grid_view<T>(I.m_i);
    } else {
        g.a.gm = grid_view<T>(I.m_i);
        g.a.gmu = grid_view<T>(I.mu_i);
    }
}

template <int ORDER, typename T>
__global__ __launch_bounds__((P2GShape<ORDER, T>::NT), (P2GShape<ORDER, T>::WAVES)) void
k_batch_p2g_brick(BrickArgs<T> g, const BatchInst* __restrict__ inst, uint32_t shift)
{
    using S = P2GShape<ORDER, T>;
    __shared__ __align__(16) T lds[S::LDS];
    __shared__ uint32_t rng[2 * 2 * S::BINS];
    __shared__ BrickItem q[2];
    for (int i = threadIdx.x; i < 4 * S::BINS; i += S::NT) rng[i] = 0u;
    P2GState<ORDER, T> st{rng, 0, false, {}};
    walk_bricks<true>(g, q, [&](const BrickItem& it, BrickWalk<true, BrickArgs<T>>& w) {
        const uint32_t k = __builtin_amdgcn_readfirstlane(it.b) >> shift;
        BrickArgs<T> gi = g;
        instance_args(inst[k], gi, false);
        p2g_one_brick<ORDER, T, false>(gi, it, w, lds, st);
    });
}

template <int ORDER, typename T, int NT, int R>
__global__ __launch_bounds__(NT) void k_batch_g2p_brick(BrickArgs<T> g,
                                                        const BatchInst* __restrict__ inst,
                                                        uint32_t shift)
{
    constexpr int FN = PB + ORDER;
    __shared__ T us[3 * FN * FN * FN];
    __shared__ BrickItem q[2];
    walk_bricks<true>(g, q, [&](const BrickItem& it, BrickWalk<true, BrickArgs<T>>& w) {
        const uint32_t k = __builtin_amdgcn_readfirstlane(it.b) >> shift;
        BrickArgs<T> gi = g;
        instance_args(inst[k], gi, true);
        g2p_one_brick<PB, ORDER, T, false, false, NT, R, true>(gi, it, w, us);
    });
}

// Particles the brick kernels left on the stale list: one work-item each,
// with its instance found by binary search.
template <int ORDER, typename T>
__global__ __launch_bounds__(256) void k_batch_p2g_stale(ApicArgs<T> a, const BatchInst* inst,
                                                         const int64_t* first, int ninst)
{
    const int64_t n = int64_t(a.list[0]);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t p = int64_t(a.perm[a.list[1 + i]]);
        const BatchInst& I = inst[instance_of(first, ninst, p)];
        ApicArgs<T> ai = a;
        ai.lat = I.lat;
        ai.gm = grid_view<T>(I.m_i);
        ai.gmu = grid_view<T>(I.mu_i);
        ApicParticle<T> pt;
        load_apic_particle<false>(ai, p, pt);
        Spline<ORDER, T> sd;
        eval_spline<ORDER, T, ORDER == 1, ORDER != 1>(ai.lat, pt.x, sd);
        if (!stencil_inside(sd, ai.gmu)) {
            atomicOr(ai.err, ERRBIT_BOUNDS);
            continue;
        }
        apic_p2g_particle<ORDER, false>(ai, pt, sd, [&](int i0, int j0, int k0, const T* mu, T m) {
            const int ni = sd.s0[0] + i0, nj = sd.s0[1] + j0, nk = sd.s0[2] + k0;
            const int64_t off = ai.gmu.off(ni, nj, nk);
#pragma unroll
            for (int d = 0; d < 3; ++d) atomic_add(ai.gmu.data + off + d * ai.gmu.cs, mu[d]);
            atomic_add(ai.gm.data + ai.gm.off(ni, nj, nk), m);
        });
    }
}

template <int ORDER, typename T>
__global__ __launch_bounds__(256) void k_batch_g2p_stale(ApicArgs<T> a, const BatchInst* inst,
                                                         const int64_t* first, int ninst)
{
    const int64_t n = int64_t(a.list[0]);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t p = int64_t(a.perm[a.list[1 + i]]);
        const BatchInst& I = inst[instance_of(first, ninst, p)];
        ApicArgs<T> ai = a;
        ai.lat = I.lat;
        ai.gmu = grid_view<T>(I.mu_i);
        ai.gnm = grid_view<T>(I.m_i);
        const int64_t pb = ai.x.base(p);
        T xp[3] = {ai.x.at(pb, 0), ai.x.at(pb, 1), ai.x.at(pb, 2)};
        Spline<ORDER, T> sd;
        eval_spline<ORDER, T, false, true>(ai.lat, xp, sd);
        if (!stencil_inside(sd, ai.gmu)) {
            atomicOr(ai.err, ERRBIT_BOUNDS);
            continue;
        }
        T cp[12];
        apic_g2p_particle<ORDER, false, false>(ai, sd, [&](int i0, int j0, int k0, T* ue) {
            const int ni = sd.s0[0] + i0, nj = sd.s0[1] + j0, nk = sd.s0[2] + k0;
            node_velocity3(ai.gmu, ai.gnm, ai.gmu.off(ni, nj, nk), ai.gnm.off(ni, nj, nk), ue);
        }, cp);
        const int64_t cb = ai.c.base(p);
#pragma unroll
        for (int r = 0; r < 12; ++r) ai.c.at(cb, r) = cp[r];
    }
}

// Zero every instance's grid windows: chunks of at most ZERO_CHUNK words,
// one workgroup per chunk at a time.
constexpr uint32_t ZERO_CHUNK = 16384;

__global__ __launch_bounds__(256) void k_batch_zero(const ZeroChunk* chunks, int64_t nchunks)
{
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (int64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const ZeroChunk ch = chunks[c];
        if (ch.wide) {
            uint4* p = reinterpret_cast<uint4*>(ch.p);
            for (uint32_t i = threadIdx.x; i < ch.n / 4; i += blockDim.x) p[i] = z;
        } else {
            for (uint32_t i = threadIdx.x; i < ch.n; i += blockDim.x) ch.p[i] = 0u;
        }
    }
}

}  // namespace pic

using namespace pic;

namespace {

// Does the window (ncomp components) cover one dense range of memory?
// Returns the element count.
bool dense_range(const pic_grid& g, int ncomp, int64_t* count)
{
    int64_t ext[4] = {g.n[0], g.n[1], g.n[2], ncomp};
    int64_t str[4] = {g.stride[0], g.stride[1], g.stride[2], ncomp > 1 ? g.cstride : 0};
    int order[4] = {0, 1, 2, 3};
    for (int a = 0; a < 4; ++a)
        for (int b = a + 1; b < 4; ++b)
            if (str[order[b]] < str[order[a]]) std::swap(order[a], order[b]);
    int64_t expect = 1;
    for (int q = 0; q < 4; ++q) {
        const int d = order[q];
        if (ext[d] == 1) continue;
        if (str[d] != expect) return false;
        expect *= ext[d];
    }
    *count = expect;
    return true;
}

bool same_window(const pic_grid& a, const pic_grid& b)
{
    for (int d = 0; d < 3; ++d)
        if (a.lo[d] != b.lo[d] || a.n[d] != b.n[d]) return false;
    return true;
}

template <typename T>
T* batch_alloc(pic_batch* b, size_t count)
{
    void* p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) return nullptr;
    b->allocs.push_back(p);
    return static_cast<T*>(p);
}

void batch_free(pic_batch* b)
{
    for (void* p : b->allocs) (void)hipFree(p);
    b->allocs.clear();
}

// The walk fields and particle fields shared by every brick of the batch.
template <typename T>
BrickArgs<T> batch_args(pic_context* ctx, const pic_batch* b, const pic_pfield* x,
                        const pic_pfield* m, const pic_pfield* c)
{
    BrickArgs<T> g{};
    g.keys = b->keys;
    g.bin_start = b->bin_start;
    g.active = b->active;
    g.items = b->items;
    g.walk = b->walk;
    g.stale = b->walk + WALK_LANES * WALK_PAD;
    g.nbricks = uint32_t(b->nb);
    g.brick_base = 0;
    g.a.np = b->np;
    g.a.perm = b->perm;
    g.a.list = nullptr;
    g.a.x = make_pfield<T>(*x);
    g.a.m = m ? make_pfield<T>(*m) : PField<T>{};
    g.a.c = make_pfield<T>(*c);
    g.a.sdim = 0;
    g.a.err = ctx->d_err;
    g.a.F = g.a.det = g.a.Finv = PField<double>{};
    g.a.dt = 0.0;
    return g;
}

// Walk items from the active list, walk counters and stale count cleared.
int batch_walk_items(pic_context* ctx, pic_batch* b)
{
    const unsigned nblk = unsigned(std::min<int64_t>((b->nb + 255) / 256, 2048));
    k_walk_items<<<nblk, 256, 0, ctx->stream>>>(b->active, b->bin_start, uint32_t(b->nb),
                                                  b->items, b->walk,
                                                  b->walk + WALK_LANES * WALK_PAD);
    return launch_status();
}

template <typename T, int ORDER>
int batch_p2g(pic_context* ctx, pic_batch* b, const pic_pfield* x, const pic_pfield* m,
              const pic_pfield* c)
{
    BrickArgs<T> g = batch_args<T>(ctx, b, x, m, c);
    if (const int rc = batch_walk_items(ctx, b)) return rc;
    {
        TimedLaunch tl(ctx, "batch_p2g");
        constexpr int NT = P2GShape<ORDER, T>::NT;
        const unsigned nbk = persistent_blocks(ctx, k_batch_p2g_brick<ORDER, T>, NT, 0, b->nb);
        k_batch_p2g_brick<ORDER, T><<<nbk, NT, 0, ctx->stream>>>(g, b->inst, b->shift);
        if (launch_status()) return PIC_ERR_HIP;
    }
    TimedLaunch tl(ctx, "batch_p2g_stale");
    g.a.list = g.stale;
    k_batch_p2g_stale<ORDER, T><<<512, 256, 0, ctx->stream>>>(g.a, b->inst, b->first, b->ninst);
    return launch_status();
}

template <typename T, int ORDER>
int batch_g2p(pic_context* ctx, pic_batch* b, const pic_pfield* x, const pic_pfield* c)
{
    BrickArgs<T> g = batch_args<T>(ctx, b, x, nullptr, c);
    if (const int rc = batch_walk_items(ctx, b)) return rc;
    {
        TimedLaunch tl(ctx, "batch_g2p");
        // 256 work-items x 3 particles unless the context asks for 128 x 4
        // (PIC_OPT_G2P_GROUP)
        if (ctx->g2p_group == 128) {
            const unsigned nbk =
                persistent_blocks(ctx, k_batch_g2p_brick<ORDER, T, 128, 4>, 128, 0, b->nb);
            k_batch_g2p_brick<ORDER, T, 128, 4><<<nbk, 128, 0, ctx->stream>>>(g, b->inst, b->shift);
        } else {
            const unsigned nbk =
                persistent_blocks(ctx, k_batch_g2p_brick<ORDER, T, BT, 3>, BT, 0, b->nb);
            k_batch_g2p_brick<ORDER, T, BT, 3><<<nbk, BT, 0, ctx->stream>>>(g, b->inst, b->shift);
        }
        if (launch_status()) return PIC_ERR_HIP;
    }
    TimedLaunch tl(ctx, "batch_g2p_stale");
    g.a.list = g.stale;
    k_batch_g2p_stale<ORDER, T><<<512, 256, 0, ctx->stream>>>(g.a, b->inst, b->first, b->ninst);
    return launch_status();
}

struct P2GPass {
    pic_context* ctx;
    pic_batch* b;
    const pic_pfield *x, *m, *c;
    template <typename T, int O>
    int operator()() const
    {
        return batch_p2g<T, O>(ctx, b, x, m, c);
    }
};

struct G2PPass {
    pic_context* ctx;
    pic_batch* b;
    const pic_pfield *x, *c;
    template <typename T, int O>
    int operator()() const
    {
        return batch_g2p<T, O>(ctx, b, x, c);
    }
};

template <typename F>
int by_order_dtype(const pic_batch* b, const F& f)
{
    if (b->dtype == PIC_F32) {
        if (b->order == 1) return f.template operator()<float, 1>();
        if (b->order == 2) return f.template operator()<float, 2>();
        return f.template operator()<float, 3>();
    }
    if (b->order == 1) return f.template operator()<double, 1>();
    if (b->order == 2) return f.template operator()<double, 2>();
    return f.template operator()<double, 3>();
}

bool valid_step_fields(const pic_batch* b, const pic_pfield* x, const pic_pfield* m,
                       const pic_pfield* c, bool need_m)
{
    if (b->np == 0) return true;
    return valid_pfield(x) && valid_pfield(c) && (!need_m || valid_pfield(m));
}

}  // namespace

extern "C" int pic_batch_create(pic_context* ctx, int order, int dtype, int ninst,
                                const pic_instance* inst, pic_batch** out)
{
    if (!out) return PIC_ERR_ARG;
    *out = nullptr;
    if (!ctx || order < 1 || order > 3 || (dtype != PIC_F32 && dtype != PIC_F64) || ninst < 1 ||
        !inst)
        return PIC_ERR_ARG;
    const size_t esz = dtype == PIC_F64 ? 8 : 4;  // bytes per grid scalar
    std::vector<BatchInst> table(static_cast<size_t>(ninst));
    std::vector<int64_t> first(static_cast<size_t>(ninst) + 1);
    std::vector<ZeroChunk> chunks;
    int64_t nb_max = 1, np = 0;
    for (int k = 0; k < ninst; ++k) {
        const pic_instance& I = inst[k];
        if (!(I.mesh.h > 0.0) || I.np < 0 || I.first != np || !valid_grid(&I.m_i) ||
            !valid_grid(&I.mu_i) || !same_window(I.m_i, I.mu_i) ||
            brick_edge(I.geom) != 4 || I.geom.brick[0] != 4)
            return PIC_ERR_ARG;
        const int64_t nb = pic_bin_count(&I.geom);
        if (nb < 0) return int(nb);
        BatchInst& d = table[size_t(k)];
        if (!aligned_knot_offset(order, PIC_NODE, I.geom.dual, d.knot_off)) return PIC_ERR_ARG;
        d.kg = make_key_geom(I.mesh, I.geom);
        d.lat = make_lattice(I.mesh, PIC_NODE);
        d.first = I.first;
        d.np = I.np;
        d.m_i = I.m_i;
        d.mu_i = I.mu_i;
        first[size_t(k)] = I.first;
        np += I.np;
        nb_max = std::max(nb_max, nb);
        // zero chunks of both windows (each must be one dense range)
        const pic_grid* w[2] = {&I.m_i, &I.mu_i};
        for (int q = 0; q < 2; ++q) {
            int64_t count = 0;
            if (!dense_range(*w[q], q == 0 ? 1 : 3, &count)) return PIC_ERR_UNSUPPORTED;
            // words: a head up to 16-byte alignment, a body of whole 16-byte
            // groups, a tail
            uint32_t* p = static_cast<uint32_t*>(w[q]->data);
            int64_t words = count * int64_t(esz / 4);
            const int64_t head = std::min<int64_t>(
                words, int64_t((16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15) / 4);
            if (head > 0) chunks.push_back(ZeroChunk{p, uint32_t(head), 0u});
            p += head;
            words -= head;
            const int64_t body = words & ~int64_t(3);
            for (int64_t o = 0; o < body; o += ZERO_CHUNK)
                chunks.push_back(
                    ZeroChunk{p + o, uint32_t(std::min<int64_t>(ZERO_CHUNK, body - o)), 1u});
            if (words > body) chunks.push_back(ZeroChunk{p + body, uint32_t(words - body), 0u});
        }
    }
    first[size_t(ninst)] = np;
    if (np >= int64_t(UINT32_MAX)) return PIC_ERR_CAPACITY;
    uint32_t shift = 0;
    while ((int64_t(1) << shift) < nb_max) ++shift;
    const int64_t nb = int64_t(ninst) << shift;
    if (nb * 64 >= int64_t(UINT32_MAX) || nb >= int64_t(INT32_MAX)) return PIC_ERR_CAPACITY;
    for (int k = 0; k < ninst; ++k) table[size_t(k)].brick_base = uint32_t(int64_t(k) << shift);

    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;
    pic_batch* b = new pic_batch();
    b->device = ctx->device;
    b->order = order;
    b->dtype = dtype;
    b->ninst = ninst;
    b->np = np;
    b->nb = nb;
    b->shift = shift;
    b->nchunks = int64_t(chunks.size());
    b->L = sort_layout(nb, np, true);
    b->inst = batch_alloc<BatchInst>(b, size_t(ninst));
    b->first = batch_alloc<int64_t>(b, size_t(ninst) + 1);
    b->chunks = batch_alloc<ZeroChunk>(b, chunks.size());
    b->keys = batch_alloc<uint32_t>(b, size_t(np));
    b->perm = batch_alloc<uint32_t>(b, size_t(np));
    b->bin_start = batch_alloc<uint32_t>(b, size_t(nb) + 1);
    b->active = batch_alloc<uint32_t>(b, size_t(nb) + 1);
    b->items = batch_alloc<BrickItem>(b, size_t(nb));
    b->walk = batch_alloc<uint32_t>(b, size_t(WALK_LANES) * WALK_PAD + 1 + size_t(np));
    b->sort_sc = batch_alloc<uint32_t>(b, b->L.bytes / 4);
    bool ok = b->inst && b->first && b->chunks && b->keys && b->perm && b->bin_start &&
              b->active && b->items && b->walk && b->sort_sc;
    ok = ok &&
         hipMemcpy(b->inst, table.data(), table.size() * sizeof(BatchInst), hipMemcpyHostToDevice) ==
             hipSuccess &&
         hipMemcpy(b->first, first.data(), first.size() * sizeof(int64_t),
                   hipMemcpyHostToDevice) == hipSuccess &&
         (chunks.empty() || hipMemcpy(b->chunks, chunks.data(), chunks.size() * sizeof(ZeroChunk),
                                      hipMemcpyHostToDevice) == hipSuccess);
    if (!ok) {
        batch_free(b);
        delete b;
        return PIC_ERR_ALLOC;
    }
    *out = b;
    return PIC_OK;
}

extern "C" int pic_batch_destroy(pic_batch* b)
{
    if (!b) return PIC_ERR_ARG;
    DeviceGuard dg(b->device);
    batch_free(b);
    delete b;
    return PIC_OK;
}

extern "C" int pic_batch_bin_sort(pic_context* ctx, pic_batch* b, const pic_pfield* x)
{
    Region region("picasso_amd::batch::bin_sort");
    if (!ctx || !b || b->device != ctx->device) return PIC_ERR_ARG;
    if (b->np > 0 && !valid_pfield(x)) return PIC_ERR_ARG;
    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;
    TimedLaunch tl(ctx, "batch_sort");
    sort_clear(ctx, b->L, b->sort_sc);
    if (b->np > 0) {
        const unsigned blk = blocks_for(b->np, SORT_PT * SORT_ITEMS);
        uint32_t* count = sort_count(b->sort_sc);
        uint32_t* unsorted = sort_unsorted(b->sort_sc, b->nb);
        if (b->dtype == PIC_F64)
            k_cell_keys<<<blk, SORT_PT, 0, ctx->stream>>>(
                b->np, BatchKeys<double>{make_pfield<double>(*x), b->inst, b->first, b->ninst, 64, 0},
                uint32_t(b->nb), b->keys, count, unsorted, ctx->d_err);
        else
            k_cell_keys<<<blk, SORT_PT, 0, ctx->stream>>>(
                b->np, BatchKeys<float>{make_pfield<float>(*x), b->inst, b->first, b->ninst, 64, 0},
                uint32_t(b->nb), b->keys, count, unsorted, ctx->d_err);
    }
    return sort_finish(ctx, b->np, b->nb, 64, 6, b->L, b->sort_sc, b->keys, b->perm, b->bin_start,
                       b->active);
}

extern "C" int pic_batch_apic_p2g(pic_context* ctx, pic_batch* b, const pic_pfield* x,
                                  const pic_pfield* m_p, const pic_pfield* c_p, uint32_t flags)
{
    Region region("picasso_amd::batch::APIC::p2g");
    if (!ctx || !b || b->device != ctx->device || !valid_step_fields(b, x, m_p, c_p, true))
        return PIC_ERR_ARG;
    if (flags & PIC_P2G_DETERMINISTIC) return PIC_ERR_UNSUPPORTED;  // single instances only
    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;
    if ((flags & PIC_P2G_ZERO_FIRST) && b->nchunks > 0) {
        TimedLaunch tl(ctx, "batch_zero");
        const unsigned blocks = unsigned(std::min<int64_t>(b->nchunks, 8192));
        k_batch_zero<<<blocks, 256, 0, ctx->stream>>>(b->chunks, b->nchunks);
        if (launch_status()) return PIC_ERR_HIP;
    }
    if (b->np == 0) return PIC_OK;
    return by_order_dtype(b, P2GPass{ctx, b, x, m_p, c_p});
}

extern "C" int pic_batch_apic_g2p_normalized(pic_context* ctx, pic_batch* b, const pic_pfield* x,
                                             const pic_pfield* c_p)
{
    Region region("picasso_amd::batch::APIC::g2p_normalized");
    if (!ctx || !b || b->device != ctx->device || !valid_step_fields(b, x, nullptr, c_p, false))
        return PIC_ERR_ARG;
    if (b->np == 0) return PIC_OK;
    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;
    return by_order_dtype(b, G2PPass{ctx, b, x, c_p});
}

extern "C" int pic_batch_apic_step(pic_context* ctx, pic_batch* b, const pic_pfield* x,
                                   const pic_pfield* m_p, const pic_pfield* c_p, uint32_t flags)
{
    Region region("picasso_amd::batch::APIC::step");
    if (const int rc = pic_batch_bin_sort(ctx, b, x)) return rc;
    if (const int rc = pic_batch_apic_p2g(ctx, b, x, m_p, c_p, flags)) return rc;
    return pic_batch_apic_g2p_normalized(ctx, b, x, c_p);
}
