// Brick machinery shared by the APIC and PolyPIC brick kernels: the walk
// over the non-empty bricks of a pic_bin_sort binning, bin ranges, the
// stale-particle list, the staged-particle layout and the node gather that
// turns per-bin stencil sums into one atomic per footprint node.
//
// Device helpers take any argument struct G with the fields of BrickWalk
// (keys, bin_start, active, walk, stale, nbricks, cell_lo, nb, knot_off).
#pragma once

#include "pic_common.hpp"
#include "pic_internal.hpp"

namespace pic {

// Binning fields every brick kernel argument struct carries.
#define PIC_BRICK_WALK_FIELDS                                                        \
    const uint32_t* keys;                                                            \
    const uint32_t* bin_start;                                                       \
    const uint32_t* active; /* [count, ids...] of non-empty bricks, or null */       \
    uint32_t* walk;         /* brick-walk counter, zero at launch */                 \
    uint32_t* stale;        /* [count, slots...] of particles for the slow path */   \
    uint32_t nbricks;                                                                \
    int32_t cell_lo[3];                                                              \
    int32_t nb[3];                                                                   \
    int32_t knot_off; /* first knot index = bin index + knot_off */

template <int BB, typename G>
__device__ __forceinline__ void brick_origin(const G& g, uint32_t b, int f0[3])
{
    const int bk = int(b % uint32_t(g.nb[2]));
    const int bj = int((b / uint32_t(g.nb[2])) % uint32_t(g.nb[1]));
    const int bi = int(b / (uint32_t(g.nb[2]) * uint32_t(g.nb[1])));
    f0[0] = g.cell_lo[0] + bi * BB + g.knot_off;
    f0[1] = g.cell_lo[1] + bj * BB + g.knot_off;
    f0[2] = g.cell_lo[2] + bk * BB + g.knot_off;
}

// Walk indices run over the active list when there is one, else over every
// brick (empty ones are skipped).
template <typename G>
__device__ __forceinline__ uint32_t walk_count(const G& g)
{
    return g.active ? g.active[0] : g.nbricks;
}

template <typename G>
__device__ __forceinline__ uint32_t walk_brick(const G& g, uint32_t t)
{
    return g.active ? g.active[1 + t] : t;
}

// Drive body(brick) over the walk.  A workgroup's first index is its
// blockIdx.  With DYN, each later one is requested from the counter when the
// previous brick starts (the request's latency hides behind that brick), so
// bricks start in list order whatever each one costs; every workgroup ends
// after drawing an index past the end.  `slot` is a two-word LDS buffer
// (alternating, so no extra barrier is needed when a brick is skipped).
// Without DYN the walk is a plain grid stride, which suits kernels whose
// bricks cost about the same (a contended counter then only adds latency).
template <bool DYN, typename G, typename Body>
__device__ __forceinline__ void walk_bricks(const G& g, uint32_t* slot, Body&& body)
{
    const uint32_t nw = walk_count(g);
    uint32_t t = blockIdx.x;
    for (int it = 0; t < nw; ++it) {
        uint32_t nxt = t + gridDim.x;
        if (DYN && threadIdx.x == 0) nxt = gridDim.x + atomicAdd(g.walk, 1u);
        const uint32_t b = walk_brick(g, t);
        if (g.bin_start[b] != g.bin_start[b + 1]) body(b);
        if constexpr (DYN) {
            if (threadIdx.x == 0) slot[it & 1] = nxt;
            __syncthreads();
            t = slot[it & 1];
        } else {
            __syncthreads();
            t = nxt;
        }
    }
}

template <typename G>
__device__ __forceinline__ void push_stale(const G& g, uint32_t s)
{
    const uint32_t at = atomicAdd(g.stale, 1u);
    g.stale[1 + at] = s;
}

// Bin ranges [rng[l], rng[BINS + l]) of brick b (slots [beg, end)) from the
// sorted keys, by neighbour comparison (empty bins: 0, 0).  Ends with a
// barrier.
template <int BINS, int NT, typename G>
__device__ __forceinline__ void bin_ranges(const G& g, uint32_t b, uint32_t beg, uint32_t end,
                                           uint32_t* rng)
{
    for (int l = threadIdx.x; l < 2 * BINS; l += NT) rng[l] = 0u;
    __syncthreads();
    const uint32_t kb = b * BINS;
    for (uint32_t s = beg + threadIdx.x; s < end; s += NT) {
        const uint32_t key = g.keys[s];
        if (s == beg || g.keys[s - 1] != key) rng[key - kb] = s;
        if (s + 1 == end || g.keys[s + 1] != key) rng[BINS + key - kb] = s + 1;
    }
    __syncthreads();
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
// brick whose first knot is f0.  Each footprint node adds the sums landing
// on it in FP64 and applies one atomic per component: component r_lo + cc
// goes to gm (cc index 0 of round 0, when gm is not null) or to component
// (r_lo + cc - c_mom) of gmu.
template <int NK, int PB, int NT, int CPR, typename T>
__device__ __forceinline__ void gather_flush(const T* sums, const int f0[3], int r_lo, int c_mom,
                                             const GView<T>& gm, const GView<T>& gmu, int* err)
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
            if (v == T(0)) continue;
            if (c < c_mom) atomic_add(gm.data + gm.off(ni, nj, nk), v);
            else atomic_add(gmu.data + off + (c - c_mom) * gmu.cs, v);
        }
    }
}

// Knot offset of the bins for (order, entity), or false when the bin
// lattice is not the spline's knot-base lattice.  In half cells: the base is
// floor(t_node - centred + order2) and the bin floor(t_node + dual).
inline bool aligned_knot_offset(int order, int entity, int dual, int* knot_off)
{
    if (entity != PIC_NODE && entity != PIC_CELL) return false;
    const int halves = (order == 2 ? 1 : 0) - (entity == PIC_CELL ? 1 : 0) - (dual ? 1 : 0);
    if (halves != 0 && halves != -2) return false;
    *knot_off = halves / 2 + (order == 1 ? 0 : -1);
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

// Fill the walk fields of g for bins; false when the brick kernels do not
// apply (then the caller falls back to the per-particle kernels).
template <typename G>
inline bool setup_walk(pic_context* ctx, int order, int entity, int64_t np, const pic_bins& bins,
                       G& g)
{
    const int bb = brick_edge(bins.geom);
    if (!bb || !bins.keys || !bins.bin_start || !bins.perm) return false;
    if (!aligned_knot_offset(order, entity, bins.geom.dual, &g.knot_off)) return false;
    const int64_t nb = pic_bin_count(&bins.geom);
    if (nb <= 0) return false;
    g.keys = bins.keys;
    g.bin_start = bins.bin_start;
    g.active = bins.active;
    g.nbricks = uint32_t(nb);
    for (int d = 0; d < 3; ++d) {
        g.cell_lo[d] = bins.geom.cell_lo[d];
        g.nb[d] = (bins.geom.ncell[d] + bb - 1) / bb;
    }
    // scratch: walk counter | stale count | stale slots[np]
    uint32_t* w = static_cast<uint32_t*>(scratch(ctx, size_t(np + 2) * sizeof(uint32_t)));
    if (!w) return false;
    g.walk = w;
    g.stale = w + 1;
    return fill_u32(ctx, w, 2, 0u) == PIC_OK;
}

}  // namespace pic
