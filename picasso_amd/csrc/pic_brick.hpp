// Brick machinery shared by the APIC and PolyPIC brick kernels: the walk
// over the non-empty bricks of a pic_bin_sort binning, chunk-local bin
// ranges, the stale-particle list, the staged-particle layout and the node
// gather that turns per-bin stencil sums into one atomic per footprint node.
//
// Device helpers take any argument struct G carrying PIC_BRICK_WALK_FIELDS.
#pragma once

#include "pic_common.hpp"
#include "pic_internal.hpp"
#include "pic_sort.hpp"

namespace pic {

// Binning fields every brick kernel argument struct carries.
#define PIC_BRICK_WALK_FIELDS                                                        \
    const uint32_t* keys;                                                            \
    const uint32_t* bin_start;                                                       \
    const uint32_t* active; /* [count, ids...] of non-empty bricks, or null */       \
    const BrickItem* items; /* one per walk index (k_walk_items or the sort's scan) */\
    uint32_t* walk;         /* WALK_LANES counters, WALK_PAD words apart, zero */    \
    uint32_t* stale;        /* [count, slots...] of particles for the slow path */   \
    uint32_t nbricks;                                                                \
    uint32_t brick_base; /* brick id of bin brick 0 (batches: the instance's first) */ \
    int32_t cell_lo[3];                                                              \
    int32_t nb[3];                                                                   \
    int32_t knot_off[3]; /* first knot index = bin index + knot_off, per dimension */

// The dynamic walk draws walk indices from WALK_LANES counters: lane x hands
// out the indices congruent to x modulo WALK_LANES, in order.  Every draw on
// one address is serialised by the memory system, so one counter shared by
// a thousand workgroups caps the brick rate; eight counters on separate
// lines keep the rate up while all workgroups still sweep one compact front
// of the list.
constexpr int WALK_LANES = 8;
constexpr int WALK_PAD = 64;  // words between counters (256 B)
constexpr uint32_t NUM_XCD = 8;  // accelerator complex dies of an MI355X, one L2 each

template <int BB, typename G>
__device__ __forceinline__ void brick_origin(const G& g, uint32_t b, int f0[3])
{
    b -= g.brick_base;
    const int bk = int(b % uint32_t(g.nb[2]));
    const int bj = int((b / uint32_t(g.nb[2])) % uint32_t(g.nb[1]));
    const int bi = int(b / (uint32_t(g.nb[2]) * uint32_t(g.nb[1])));
    f0[0] = g.cell_lo[0] + bi * BB + g.knot_off[0];
    f0[1] = g.cell_lo[1] + bj * BB + g.knot_off[1];
    f0[2] = g.cell_lo[2] + bk * BB + g.knot_off[2];
}

// Walk indices run over the active list when there is one, else over every
// brick (empty ones are skipped).
template <typename G>
__device__ __forceinline__ uint32_t walk_count(const G& g)
{
    return g.active ? g.active[0] : g.nbricks;
}

// Item of walk index t (past the end: t with an empty range).
template <typename G>
__device__ __forceinline__ BrickItem load_item(const G& g, uint32_t t, uint32_t nw)
{
    if (t < nw) return g.items[t];
    return BrickItem{t, 0u, 0u, 0u};
}

// The items of the walk in one flat array, so that a workgroup resolves its
// next brick with one load instead of the chain active -> bin_start.  Also
// clears the walk counters and the stale count.
template <int = 0>
__global__ __launch_bounds__(256) void k_walk_items(const uint32_t* active,
                                                    const uint32_t* bin_start, uint32_t nbricks,
                                                    BrickItem* items, uint32_t* walk,
                                                    uint32_t* stale)
{
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < WALK_LANES * WALK_PAD; i += blockDim.x) walk[i] = 0u;
        if (threadIdx.x == 0) stale[0] = 0u;
    }
    const uint32_t nw = active ? active[0] : nbricks;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nw; t += gridDim.x * blockDim.x) {
        const uint32_t b = active ? active[1 + t] : t;
        items[t] = BrickItem{t, b, bin_start[b], bin_start[b + 1]};
    }
}

// Walk state of a workgroup.  Brick n's item sits in q[n & 1] (LDS).
// Work-item 0 draws the walk index of brick n+1 at brick n's start
// (draw), loads its item once the draw has returned (fetch) and stores it
// into q[(n+1) & 1] (publish); after the next barrier every work-item can
// read it (peek), e.g. to prefetch the next brick's slots.  The bodies call
// fetch() and publish() once per brick, in that order, at points where
// their latency hides behind the brick's own loads.
template <bool DYN, typename G>
struct BrickWalk {
    const G& g;
    BrickItem* q;
    uint32_t nw;
    uint32_t lane, nl;  // counter lane of this workgroup, lanes in use
    int n;
    uint32_t t_next;  // work-item 0
    BrickItem nxt;    // work-item 0

    __device__ __forceinline__ void draw(uint32_t t)
    {
        if (threadIdx.x != 0) return;
        if constexpr (DYN)
            t_next = lane + nl * (gridDim.x / nl + atomicAdd(g.walk + lane * WALK_PAD, 1u));
        else
            t_next = t + gridDim.x;
    }
    __device__ __forceinline__ void fetch()
    {
        if (threadIdx.x == 0) nxt = load_item(g, t_next, nw);
    }
    __device__ __forceinline__ void publish()
    {
        if (threadIdx.x == 0) q[(n + 1) & 1] = nxt;
    }
    __device__ __forceinline__ BrickItem peek() const { return q[(n + 1) & 1]; }
    __device__ __forceinline__ bool has_work(const BrickItem& it) const
    {
        return it.t < nw && it.beg < it.end;
    }
};

// Drive body(item, walk) over the walk.  Without DYN the walk is a plain
// grid stride (workgroup b takes b, b + grid, ...; in the XCD-contiguous
// order below) that needs no counters: the colour passes use it, one launch
// per colour with no clear in between.  With DYN, when the grid
// is a multiple of WALK_LANES, workgroup b (b = 8m + r) starts at walk index
// 8m + lane with lane = (r + m) mod 8 and then draws from counter `lane`;
// so each lane's workgroups sit on every XCD (blockIdx modulo 8 picks the
// XCD), the lanes advance at the same pace and bricks start in list order
// whatever each one costs: neighbouring bricks run at the same time and
// share their halo nodes in L2.  Otherwise one counter serves the grid.
// Every workgroup ends after drawing an index past the end.  q is a
// two-item LDS buffer.  Bodies do not end with a barrier: the walk places
// one after each brick.
template <bool DYN, typename G, typename Body>
__device__ __forceinline__ void walk_bricks(const G& g, BrickItem* q, Body&& body)
{
    uint32_t nl = 1, lane = 0, t0 = blockIdx.x;
    if (DYN && gridDim.x % WALK_LANES == 0) {
        const uint32_t m = blockIdx.x / WALK_LANES;
        nl = WALK_LANES;
        lane = (blockIdx.x + m) % WALK_LANES;
        t0 = lane + WALK_LANES * m;
    } else if (!DYN && gridDim.x % NUM_XCD == 0) {
        // Static walk: workgroups are dealt to the XCDs round-robin by
        // blockIdx, so give each XCD a contiguous run of gridDim / 8 walk
        // indices per round; neighbouring bricks then share their halo
        // rows in one XCD's L2 instead of fetching them into eight.
        t0 = (blockIdx.x % NUM_XCD) * (gridDim.x / NUM_XCD) + blockIdx.x / NUM_XCD;
    }
    BrickWalk<DYN, G> w{g, q, walk_count(g), lane, nl, 0, 0u, BrickItem{}};
    if (threadIdx.x == 0) q[0] = load_item(g, t0, w.nw);
    __syncthreads();
    for (;; ++w.n) {
        const BrickItem cur = q[w.n & 1];
        if (cur.t >= w.nw) break;
        w.draw(cur.t);
        if (cur.beg < cur.end) {
            body(cur, w);
        } else {
            w.fetch();
            w.publish();
        }
        __syncthreads();
    }
}

// Sorted slots of one staging chunk [c0, c1), K per work-item (slot
// c0 + threadIdx.x + k NT): the particle index and the bin keys of the slot
// and its neighbours inside the chunk (~0 past either end), from which the
// chunk-local bin ranges follow.  Loaded either while the previous brick
// is flushed or at staging time.
template <int K, int NT>
struct ChunkSlots {
    uint32_t perm[K], key[K], prv[K], nxt[K];

    template <typename G>
    __device__ __forceinline__ void load(const G& g, const uint32_t* perm_v, uint32_t c0,
                                         uint32_t c1)
    {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t s = c0 + threadIdx.x + uint32_t(k * NT);
            if (s < c1) {
                perm[k] = perm_v[s];
                key[k] = g.keys[s];
                prv[k] = s > c0 ? g.keys[s - 1] : ~0u;
                nxt[k] = s + 1 < c1 ? g.keys[s + 1] : ~0u;
            }
        }
    }
    // Chunk-local range [rng[bin], rng[BINS + bin]) of slot k's bin (chunk
    // position q), from the neighbour keys.
    template <int BINS>
    __device__ __forceinline__ void note(int k, int q, uint32_t kb, uint32_t* rng) const
    {
        if (prv[k] != key[k]) rng[key[k] - kb] = uint32_t(q);
        if (nxt[k] != key[k]) rng[BINS + key[k] - kb] = uint32_t(q + 1);
    }
};

// Per-workgroup staging state of a p2g brick walk: double-buffered
// chunk-local bin ranges (the buffer of the next chunk is cleared while the
// current one is staged) and the slots of the next brick's first chunk,
// prefetched while the current brick is flushed.
template <int K, int NT, int BINS>
struct BrickStage {
    static_assert(2 * BINS <= NT, "range buffers are cleared one word per work-item");
    uint32_t* rng;  // LDS [2][2 * BINS], zero at the start of the walk
    int par;        // range buffer of the next chunk
    bool have;      // cs holds the first chunk of the current brick
    ChunkSlots<K, NT> cs;

    // Slots of chunk [c0, c1) of a brick starting at beg; returns the range
    // buffer the chunk's notes go to.
    template <typename G>
    __device__ __forceinline__ uint32_t* begin_chunk(const G& g, const uint32_t* perm_v,
                                                     uint32_t beg, uint32_t c0, uint32_t c1)
    {
        if (!(have && c0 == beg)) cs.load(g, perm_v, c0, c1);
        return rng + par * (2 * BINS);
    }
    // Before the barrier that ends the staging of a chunk.
    __device__ __forceinline__ void end_chunk()
    {
        if (threadIdx.x < 2 * BINS) rng[(par ^ 1) * (2 * BINS) + threadIdx.x] = 0u;
        par ^= 1;
    }
    // After the barrier that follows the walk's publish(): prefetch the
    // next brick's first chunk.
    template <typename G, typename W>
    __device__ __forceinline__ void prefetch_next(const G& g, const uint32_t* perm_v, const W& w,
                                                  uint32_t cap)
    {
        const BrickItem nx = w.peek();
        have = w.has_work(nx);
        if (have) cs.load(g, perm_v, nx.beg, min(nx.end, nx.beg + cap));
    }
};

// Direct copy of one 4-byte scalar per active work-item from global memory
// into LDS, without a register: lane i of the wave writes wave_dst + 4 i
// (wave_dst must be the same for the whole wave; lanes masked off write
// nothing).  Completion is counted by the vector-memory counter, so readers
// in other waves need a vmcnt(0) wait before the barrier they sync on.
__device__ __forceinline__ void load_to_lds(const void* src, void* wave_dst)
{
    __builtin_amdgcn_global_load_lds(
        (const void __attribute__((address_space(1)))*)src,
        (void __attribute__((address_space(3)))*)wave_dst, 4, 0, 0);
}

// s_waitcnt operand for vmcnt(0), other counters left free (gfx9 layout:
// vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
constexpr int WAIT_VMCNT0 = 0x0F70;
// s_waitcnt operand for lgkmcnt(0), vmcnt and expcnt left free.
constexpr int WAIT_LGKMCNT0 = 0xC07F;

template <typename G>
__device__ __forceinline__ void push_stale(const G& g, uint32_t s)
{
    const uint32_t at = atomicAdd(g.stale, 1u);
    g.stale[1 + at] = s;
}

constexpr double STALE_F = 4.0;  // staged f0 of a particle the brick skips

// The components of one bin-sum slot, read with one wide LDS load.
template <typename T, int N>
struct alignas(sizeof(T) * N) SumSlot {
    T c[N];
};

// Scalar offset of group g of staged particle q.
__device__ __forceinline__ int stage_at(int q, int g)
{
    return q * 16 + 4 * (g ^ ((q >> 2) & 3));
}

// Node gather of one round of bin sums: sums[slot * CPR + c] with slot =
// offset * BINS + bin, offsets (oi, oj, ok) in [0, NK)^3 and bins of a PB^3
// brick whose first knot is f0.  Each footprint node adds the (at most NK^3)
// sums landing on it in FP64, which stands in for compensated FP32
// summation at less cost than a two-sum (an FP64 add issues at the FP32
// rate), and applies one atomic per component: component r_lo + cc
// goes to gm (cc index 0 of round 0, when gm is not null) or to component
// (r_lo + cc - c_mom) of gmu; components from nwrite on are not written.
// PLAIN replaces the atomics by plain read-modify-writes, for walks in
// which no other brick touches the footprint at the same time.
template <int NK, int PB, int NT, int CPR, typename T, bool PLAIN = false>
__device__ __forceinline__ void gather_flush(const T* sums, const int f0[3], int r_lo, int c_mom,
                                             const GView<T>& gm, const GView<T>& gmu, int* err,
                                             int nwrite = 1 << 30)
{
    constexpr int FN = PB + NK - 1;
    constexpr int FV = FN * FN * FN;
    constexpr int BINS = PB * PB * PB;
    for (int q = threadIdx.x; q < FV; q += NT) {
        const int fi = q / (FN * FN), fj = (q / FN) % FN, fk = q % FN;
        double tot[CPR];
#pragma unroll
        for (int cc = 0; cc < CPR; ++cc) tot[cc] = 0.0;
#pragma unroll 1
        for (int oi = 0; oi < NK; ++oi)
#pragma unroll
            for (int oj = 0; oj < NK; ++oj)
#pragma unroll
                for (int ok = 0; ok < NK; ++ok) {
                    const int bi = fi - oi, bj = fj - oj, bk = fk - ok;
                    if (bi < 0 || bi >= PB || bj < 0 || bj >= PB || bk < 0 || bk >= PB) continue;
                    const int slot = ((oi * NK + oj) * NK + ok) * BINS + (bi * PB + bj) * PB + bk;
                    const SumSlot<T, CPR> v = *reinterpret_cast<const SumSlot<T, CPR>*>(sums + slot * CPR);
#pragma unroll
                    for (int cc = 0; cc < CPR; ++cc) tot[cc] += double(v.c[cc]);
                }
        bool any = false;
#pragma unroll
        for (int cc = 0; cc < CPR; ++cc) any = any || tot[cc] != 0.0;
        if (!any) continue;
        const int ni = f0[0] + fi, nj = f0[1] + fj, nk = f0[2] + fk;
        if (!(gmu.covers(0, ni, ni) && gmu.covers(1, nj, nj) && gmu.covers(2, nk, nk))) {
            atomicOr(err, ERRBIT_BOUNDS);
            continue;
        }
        const int64_t off = gmu.off(ni, nj, nk);
#pragma unroll
        for (int cc = 0; cc < CPR; ++cc) {
            const int c = r_lo + cc;
            const T v = T(tot[cc]);
            if (v == T(0) || c >= nwrite) continue;
            T* dst = c < c_mom ? gm.data + gm.off(ni, nj, nk) : gmu.data + off + (c - c_mom) * gmu.cs;
            if constexpr (PLAIN) *dst = *dst + v;
            else atomic_add(dst, v);
        }
    }
}

// Colour of brick b (0..7): the parities of its brick coordinates.  Two
// bricks of one colour are at least 2 bricks apart in some dimension, so
// their node footprints (at most 2 PB nodes wide) never overlap, and a walk
// over one colour's bricks can flush its nodes without atomics.
template <typename G>
__device__ __forceinline__ uint32_t brick_colour(const G& g, uint32_t b)
{
    b -= g.brick_base;
    const uint32_t bk = b % uint32_t(g.nb[2]);
    const uint32_t bj = (b / uint32_t(g.nb[2])) % uint32_t(g.nb[1]);
    const uint32_t bi = b / (uint32_t(g.nb[2]) * uint32_t(g.nb[1]));
    return ((bi & 1u) << 2) | ((bj & 1u) << 1) | (bk & 1u);
}

// The walk items grouped by brick colour, for the deterministic p2g: one
// workgroup counts the colours, scans the eight counts and scatters the
// items, renumbering each item's walk index within its colour.
// colour[0..7] receives the counts, colour[8..15] the first item of each
// colour in `out`.  (The order inside a colour follows the LDS atomics and
// does not matter: one colour's bricks touch disjoint nodes.)
template <typename G>
__global__ __launch_bounds__(1024) void k_colour_items(G g, BrickItem* out, uint32_t* colour)
{
    __shared__ uint32_t cnt[8], cur[8];
    if (threadIdx.x < 8) cnt[threadIdx.x] = cur[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t nw = walk_count(g);
    for (uint32_t t = threadIdx.x; t < nw; t += blockDim.x)
        atomicAdd(&cnt[brick_colour(g, g.items[t].b)], 1u);
    __syncthreads();
    if (threadIdx.x < 8) {
        uint32_t off = 0;
        for (uint32_t c = 0; c < threadIdx.x; ++c) off += cnt[c];
        colour[threadIdx.x] = cnt[threadIdx.x];
        colour[8 + threadIdx.x] = off;
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < nw; t += blockDim.x) {
        BrickItem it = g.items[t];
        const uint32_t c = brick_colour(g, it.b);
        const uint32_t slot = atomicAdd(&cur[c], 1u);
        it.t = slot;
        out[colour[8 + c] + slot] = it;
    }
}

// Knot offset of the bins for (order, entity) per dimension, or false when
// the bin lattice is not the spline's knot-base lattice in every dimension.
// In half cells: the base is floor(t_node - centred + order2) and the bin
// floor(t_node + dual), with centred and dual taken per dimension (staggered
// Face/Edge lattices are node-like in some dimensions, cell-like in the
// others).
inline bool aligned_knot_offset(int order, int entity, int dual, int32_t knot_off[3])
{
    const int mask = dual_mask(dual);
    if (mask < 0) return false;
    for (int d = 0; d < 3; ++d) {
        const int halves =
            (order == 2 ? 1 : 0) - (cell_centred(entity, d) ? 1 : 0) - ((mask >> d) & 1);
        if (halves != 0 && halves != -2) return false;
        knot_off[d] = halves / 2 + (order == 1 ? 0 : -1);
    }
    return true;
}

// Cubic brick edge of the binning, or 0 when the brick kernels do not
// support it.
inline int brick_edge(const pic_binning& geom)
{
    const int e = geom.brick[0];
    if (geom.brick[1] != e || geom.brick[2] != e) return 0;
    return (e == 4 || e == 8) ? e : 0;
}

// The walk fields of g that follow from the bins (all but the scratch
// pointers items, walk and stale).  PIC_ERR_UNSUPPORTED when the brick
// kernels do not apply (the caller then falls back to the per-particle
// kernels).
template <typename G>
inline int walk_geometry(int order, int entity, const pic_bins& bins, G& g)
{
    const int bb = brick_edge(bins.geom);
    if (!bb || !bins.keys || !bins.bin_start || !bins.perm) return PIC_ERR_UNSUPPORTED;
    if (!aligned_knot_offset(order, entity, bins.geom.dual, g.knot_off))
        return PIC_ERR_UNSUPPORTED;
    const int64_t nb = pic_bin_count(&bins.geom);
    if (nb <= 0) return PIC_ERR_UNSUPPORTED;
    g.keys = bins.keys;
    g.bin_start = bins.bin_start;
    g.active = bins.active;
    g.nbricks = uint32_t(nb);
    g.brick_base = 0;
    for (int d = 0; d < 3; ++d) {
        g.cell_lo[d] = bins.geom.cell_lo[d];
        g.nb[d] = (bins.geom.ncell[d] + bb - 1) / bb;
    }
    return PIC_OK;
}

// Fill the walk fields of g for bins and list the walk items in scratch.
// PIC_ERR_UNSUPPORTED when the brick kernels do not apply, PIC_ERR_ALLOC
// when the scratch cannot grow (inside a capture, or while a step graph of
// the context is alive).
template <typename G>
inline int setup_walk(pic_context* ctx, int order, int entity, int64_t np, const pic_bins& bins,
                      G& g, size_t extra_bytes = 0, void** extra = nullptr)
{
    if (const int rc = walk_geometry(order, entity, bins, g)) return rc;
    const int64_t nb = int64_t(g.nbricks);
    // scratch: items[nbricks] | walk counters | stale count | stale slots[np]
    // | the caller's extra bytes (16-byte aligned)
    const size_t ibytes = size_t(nb) * sizeof(BrickItem);
    const size_t words = size_t(WALK_LANES) * WALK_PAD + 1 + size_t(np);
    const size_t wbytes = (words * sizeof(uint32_t) + 15) & ~size_t(15);
    char* base = static_cast<char*>(scratch(ctx, ibytes + wbytes + extra_bytes));
    if (!base) return PIC_ERR_ALLOC;
    if (extra) *extra = base + ibytes + wbytes;
    BrickItem* items = reinterpret_cast<BrickItem*>(base);
    uint32_t* w = reinterpret_cast<uint32_t*>(base + ibytes);
    g.items = items;
    g.walk = w;
    g.stale = w + WALK_LANES * WALK_PAD;
    const unsigned nblk = unsigned(std::min<int64_t>((nb + 255) / 256, 2048));
    k_walk_items<<<nblk, 256, 0, ctx->stream>>>(bins.active, bins.bin_start, uint32_t(nb), items,
                                                  w, g.stale);
    return launch_status();
}

}  // namespace pic
