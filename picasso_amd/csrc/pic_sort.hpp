// Bin geometry and key computation shared by the sort and the brick kernels.
#pragma once

#include "pic_common.hpp"

namespace pic {

struct KeyGeom {
    double low[3];
    double rdx;
    int32_t cell_lo[3];
    int32_t ncell[3];
    int32_t brick[3];
    int32_t nb[3];
    uint32_t bvol;
    int32_t dual;  // bit d: dimension d bins by node-centred cell
    // log2 of the brick edges and volume when all are powers of two (the
    // usual case), so the brick / bin split is shifts and masks; else -1.
    int32_t bshift[3];
    int32_t vshift;
};

KeyGeom make_key_geom(const pic_mesh& mesh, const pic_binning& b);

// Key of particle p (0xffffffff when outside the bins); *brick receives the
// brick index.
template <typename T>
__device__ __forceinline__ uint32_t cell_key(const KeyGeom& g, const PField<T>& x, int64_t p,
                                             uint32_t* brick)
{
    const int64_t xb = x.base(p);
    int32_t rel[3];
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double c = bin_coord((double)x.at(xb, d), g.low[d], g.rdx, (g.dual >> d) & 1) -
                         (double)g.cell_lo[d];
        ok = ok && (c >= 0.0) && (c < (double)g.ncell[d]);
        rel[d] = ok ? (int32_t)c : 0;
    }
    if (!ok) {
        *brick = 0;
        return 0xffffffffu;
    }
    uint32_t bi, bj, bk, li, lj, lk;
    if (g.vshift >= 0) {
        bi = uint32_t(rel[0]) >> g.bshift[0];
        bj = uint32_t(rel[1]) >> g.bshift[1];
        bk = uint32_t(rel[2]) >> g.bshift[2];
        li = uint32_t(rel[0]) & uint32_t(g.brick[0] - 1);
        lj = uint32_t(rel[1]) & uint32_t(g.brick[1] - 1);
        lk = uint32_t(rel[2]) & uint32_t(g.brick[2] - 1);
    } else {
        bi = rel[0] / g.brick[0];
        bj = rel[1] / g.brick[1];
        bk = rel[2] / g.brick[2];
        li = rel[0] % g.brick[0];
        lj = rel[1] % g.brick[1];
        lk = rel[2] % g.brick[2];
    }
    const uint32_t b = (bi * g.nb[1] + bj) * g.nb[2] + bk;
    *brick = b;
    return b * g.bvol + (li * g.brick[1] + lj) * g.brick[2] + lk;
}

// ---------------------------------------------------------------------------
// The sort's passes for callers with a key pass of their own (batches of
// instances): sort_layout sizes the scratch, sort_clear zeroes its first
// nzero words, the caller runs k_cell_keys with its key functor on the
// layout's count and unsorted words, and sort_finish runs the scans, the
// scatter and the per-brick pass.

// Scratch of one sort: count[nb+1] | cursor[nb+1] | unsorted flag (2 words)
// | tile sums (64-bit) | act[nb+1] | tmp[np] | tmpk[np] | big[np], in
// 4-byte words.  act holds the active-brick list when the caller does not
// ask for it.  The first nzero words must start at zero.
struct SortLayout {
    int64_t ntile, nzero, head;
    size_t bytes;
};

SortLayout sort_layout(int64_t nb, int64_t np, bool caller_active);
inline uint32_t* sort_count(uint32_t* sc) { return sc; }
inline uint32_t* sort_unsorted(uint32_t* sc, int64_t nb) { return sc + 2 * (nb + 1); }
void sort_clear(pic_context* ctx, const SortLayout& L, uint32_t* sc);
int sort_finish(pic_context* ctx, int64_t np, int64_t nb, uint32_t bvol, int32_t vshift,
                const SortLayout& L, uint32_t* sc, uint32_t* keys, uint32_t* perm,
                uint32_t* bin_start, uint32_t* active);

// Wave-aggregated count: atomicAdd of 1 per active lane on base[key], lanes
// with equal keys sharing one atomic.  Nothing is returned, so the atomics
// are fire-and-forget and a work-item's later items do not wait for them.
__device__ __forceinline__ void wave_agg_count(uint32_t* base, uint32_t key, bool active)
{
    const int lane = __lane_id();
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lkey = __shfl(key, leader);
        const uint64_t same = __ballot(active && key == lkey) & pending;
        if (lane == leader) (void)atomicAdd(base + lkey, (uint32_t)__popcll(same));
        pending &= ~same;
    }
}

// Particles per work-item in the per-particle sort passes (the grid is a
// quarter of the particles).  The key pass waits for each item's position
// before loading the next; at 40 registers, eight waves per SIMD hide that
// (issuing all twelve loads first measured no faster).
constexpr int SORT_ITEMS = 4;
constexpr int SORT_PT = 256;  // work-items per block of those passes

// Keys of one mesh's particles: kf(p, &brick) is cell_key.
template <typename T>
struct MeshKeys {
    PField<T> x;
    KeyGeom g;
    __device__ __forceinline__ void begin(int64_t) {}
    __device__ __forceinline__ uint32_t operator()(int64_t p, uint32_t* brick) const
    {
        return cell_key(g, x, p, brick);
    }
};

// Key pass: kf(p, &brick) gives particle p's key (0xffffffff outside the
// bins) and its brick, after kf.begin(first particle of the workgroup) has
// run once per workgroup.  Particles outside the bins are counted in
// count[nbricks] (they are placed after every brick, so perm stays a
// permutation even then).  Keys go to
// keys[] in input order.  *unsorted is raised (by one atomic per
// workgroup) when some particle's key is smaller than its predecessor's;
// the predecessor of a workgroup's first particle is recomputed from its
// position.
template <typename KeyFn>
__global__ __launch_bounds__(SORT_PT) void k_cell_keys(int64_t np, KeyFn kf, uint32_t nbricks,
                                                       uint32_t* keys, uint32_t* count,
                                                       uint32_t* unsorted, int* err)
{
    __shared__ uint32_t skey[SORT_PT * SORT_ITEMS];
    const int64_t blk0 = int64_t(blockIdx.x) * (SORT_PT * SORT_ITEMS);
    const int64_t p0 = blk0 + threadIdx.x;
    kf.begin(blk0);
    uint32_t key[SORT_ITEMS], b[SORT_ITEMS];
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
        const int64_t p = p0 + k * SORT_PT;
        key[k] = 0xffffffffu;
        b[k] = 0;
        if (p < np) key[k] = kf(p, &b[k]);
    }
    uint32_t before = 0u;  // key of the workgroup's predecessor particle
    if (threadIdx.x == 0 && blk0 > 0) {
        uint32_t bb;
        before = kf(blk0 - 1, &bb);
    }
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
        const int64_t p = p0 + k * SORT_PT;
        const bool live = p < np;
        skey[k * SORT_PT + threadIdx.x] = key[k];
        if (live) {
            keys[p] = key[k];
            if (key[k] == 0xffffffffu) {
                atomicOr(err, ERRBIT_CELL);
                b[k] = nbricks;
            }
        }
        wave_agg_count(count, b[k], live);
    }
    __syncthreads();
    bool descent = false;
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
        const int q = k * SORT_PT + int(threadIdx.x);
        if (blk0 + q >= np) continue;
        const uint32_t prev = q > 0 ? skey[q - 1] : before;
        descent = descent || key[k] < prev;
    }
    if (__syncthreads_or(descent) && threadIdx.x == 0) atomicOr(unsorted, 1u);
}

}  // namespace pic
