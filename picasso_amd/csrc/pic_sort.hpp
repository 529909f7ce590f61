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

// One brick of a walk over the non-empty bricks: walk index t, brick id b
// and its slot range [beg, end) in the sorted order.
struct alignas(16) BrickItem {
    uint32_t t, b, beg, end;
};

// Byte ranges that a key pass clears beside its keys, for a caller that
// would otherwise clear them with launches of its own (pic_apic_step clears
// the grid windows and the brick walk's counters this way, so the clearing
// overlaps the key pass instead of following it).  Range r is bytes
// [0, bytes[r]) from base[r]; its 16-byte words (base 16-byte aligned) are
// shared out over `blocks` extra workgroups interleaved with the key
// workgroups, `per` words each, and the last bytes[r] % 16 bytes by the
// first of those workgroups.
struct ZeroJob {
    static constexpr int MAXR = 6;
    static constexpr int64_t PER = 1024;  // 16-byte words per workgroup (16 KB)
    char* base[MAXR];
    int64_t bytes[MAXR];
    int nr = 0;
    uint32_t blocks = 0;
    int64_t words = 0;
    int64_t per = PER;

    // false when full or when base is not 16-byte aligned for its words
    bool add(void* p, int64_t nbytes)
    {
        if (nr == MAXR || (nbytes >= 16 && (reinterpret_cast<uintptr_t>(p) & 15) != 0))
            return false;
        base[nr] = static_cast<char*>(p);
        bytes[nr] = nbytes;
        ++nr;
        words += nbytes >> 4;
        blocks = uint32_t((words + PER - 1) / PER);
        if (blocks == 0) blocks = 1;  // the tails alone
        return true;
    }
};

// Clearing workgroup zi of a ZeroJob: words [zi per, (zi + 1) per) of the
// ranges laid end to end, and the tails when zi is 0.
__device__ __forceinline__ void zero_share(const ZeroJob& z, uint32_t zi)
{
    const int64_t w0 = int64_t(zi) * z.per, w1 = w0 + z.per;
    const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
    int64_t r0 = 0;
    for (int r = 0; r < z.nr; ++r) {
        const int64_t n = z.bytes[r] >> 4;
        const int64_t a = w0 > r0 ? w0 : r0, b = w1 < r0 + n ? w1 : r0 + n;
        uint4* p = reinterpret_cast<uint4*>(z.base[r]);
        for (int64_t w = a + threadIdx.x; w < b; w += blockDim.x) p[w - r0] = zero;
        if (zi == 0 && int64_t(threadIdx.x) < (z.bytes[r] & 15))
            z.base[r][(n << 4) + threadIdx.x] = 0;
        r0 += n;
    }
}

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
// items (may be null): the walk items of the non-empty bricks, written by
// the scan for brick kernels that would otherwise list them in a prologue.
int sort_finish(pic_context* ctx, int64_t np, int64_t nb, uint32_t bvol, int32_t vshift,
                const SortLayout& L, uint32_t* sc, uint32_t* keys, uint32_t* perm,
                uint32_t* bin_start, uint32_t* active, BrickItem* items = nullptr);
// The active-brick list the sort keeps in its scratch when the caller
// passes none.
uint32_t* sort_active(const SortLayout& L, uint32_t* sc);
// The whole sort of a mesh's particles (pic_bin_sort) with a clearing job
// run beside its key pass and, when items is not null, the walk items.
int bin_sort_with(pic_context* ctx, int dtype, const pic_mesh& mesh, int64_t np,
                  const pic_pfield* x, const pic_binning& geom, int64_t nb, const SortLayout& L,
                  uint32_t* sc, uint32_t* keys, uint32_t* perm, uint32_t* bin_start,
                  uint32_t* active, const ZeroJob& z, BrickItem* items);


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

// Key-pass workgroups for np particles (a clearing job adds its own).
inline unsigned key_blocks(int64_t np)
{
    return unsigned((np + SORT_PT * SORT_ITEMS - 1) / (SORT_PT * SORT_ITEMS));
}

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
// position.  blk0: the workgroup's first particle.
template <typename KeyFn>
__device__ __forceinline__ void cell_keys_block(int64_t blk0, int64_t np, KeyFn& kf,
                                                uint32_t nbricks, uint32_t* keys, uint32_t* count,
                                                uint32_t* unsorted, int* err, uint32_t* skey)
{
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

template <typename KeyFn>
__global__ __launch_bounds__(SORT_PT) void k_cell_keys(int64_t np, KeyFn kf, uint32_t nbricks,
                                                       uint32_t* keys, uint32_t* count,
                                                       uint32_t* unsorted, int* err)
{
    __shared__ uint32_t skey[SORT_PT * SORT_ITEMS];
    cell_keys_block(int64_t(blockIdx.x) * (SORT_PT * SORT_ITEMS), np, kf, nbricks, keys, count,
                    unsorted, err, skey);
}

// The key pass with a clearing job: the grid holds z.blocks more
// workgroups, and workgroup w clears when floor((w + 1) Z / W) exceeds
// floor(w Z / W) (W workgroups, Z of them clearing), which spreads the
// clearing evenly over the key pass; the others take the key workgroups in
// order.  A kernel of its own, so that the plain key pass keeps its
// workgroup index on the scalar unit (a batch's per-workgroup instance
// search runs on scalar loads there).
template <typename KeyFn>
__global__ __launch_bounds__(SORT_PT) void k_cell_keys_clear(int64_t np, KeyFn kf,
                                                             uint32_t nbricks, uint32_t* keys,
                                                             uint32_t* count, uint32_t* unsorted,
                                                             int* err, ZeroJob z)
{
    __shared__ uint32_t skey[SORT_PT * SORT_ITEMS];
    const uint64_t W = gridDim.x, Z = z.blocks;
    const uint64_t before = uint64_t(blockIdx.x) * Z / W;
    if ((uint64_t(blockIdx.x) + 1) * Z / W > before) {
        zero_share(z, uint32_t(before));
        return;
    }
    const uint32_t kblk = __builtin_amdgcn_readfirstlane(blockIdx.x - uint32_t(before));
    cell_keys_block(int64_t(kblk) * (SORT_PT * SORT_ITEMS), np, kf, nbricks, keys, count,
                    unsorted, err, skey);
}

}  // namespace pic
