// Cell-binning particle sort (no reference counterpart: Picasso never sorts
// particles).  Stable counting sort by bin key:
//   key = brick_linear * brick_volume + bin_in_brick_linear (both i-major),
// so a brick's particles are contiguous and, inside it, ordered by bin and
// then by input index.
//
// Passes
//  (1) keys in input order, written to the caller's keys array, per-brick
//      counts (wave-aggregated atomics), and a flag raised when some key is
//      smaller than its predecessor's;
//  (2) exclusive scan of the counts into bin_start, and the list of
//      non-empty bricks, in two passes over tiles of counts;
//  (3) flag down (keys already non-decreasing in input order, as when the
//      particles are stored in bin order): the stable sort is the identity,
//      perm[s] = s and the keys are already in place; passes 3-4 do nothing
//      more.  Flag up: scatter of particle indices and their keys into their
//      brick segments (wave-aggregated cursors: order inside a segment
//      depends on scheduling);
//  (4) one workgroup per brick: LDS counting sort by bin, then each bin's
//      indices sorted ascending, which restores the stable order bit-exactly
//      whatever order pass 3 produced.  Writes perm and the sorted keys.
#include <mutex>

#include "pic_common.hpp"
#include "pic_internal.hpp"
#include "pic_sort.hpp"

namespace pic {

constexpr int SORT_LDS_CAP = 2048;     // particles per brick placed in LDS
constexpr int SORT_BT = 256;           // work-items per brick-sort workgroup
constexpr int SORT_MAX_BVOL = 4096;    // bins per brick
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

// Wave-aggregated slot claim in two phases, for several calls in flight:
// issue() does the aggregated atomics and keeps, per lane, its group's
// leader and its rank (among equal-key lanes, in lane order); resolve()
// reads the leader's returned base.  The atomics of all
// issued calls are then outstanding together instead of one at a time.
struct AggTicket {
    uint32_t base;  // returned old value (meaningful in the leader lane)
    int leader;
    uint32_t rank;
};

__device__ __forceinline__ AggTicket wave_agg_issue(uint32_t* base, uint32_t key, bool active)
{
    const int lane = __lane_id();
    uint64_t pending = __ballot(active);
    AggTicket t{0u, 0, 0u};
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lkey = __shfl(key, leader);
        const uint64_t same = __ballot(active && key == lkey) & pending;
        if (lane == leader) t.base = atomicAdd(base + lkey, (uint32_t)__popcll(same));
        if ((same >> lane) & 1ull) {
            t.leader = leader;
            t.rank = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        }
        pending &= ~same;
    }
    return t;
}

__device__ __forceinline__ uint32_t wave_agg_resolve(const AggTicket& t)
{
    return uint32_t(__shfl((int)t.base, t.leader)) + t.rank;
}

// Particles per work-item and work-items per block of the per-particle
// passes (SORT_ITEMS, SORT_PT) are in pic_sort.hpp with the key pass.

// ---------------------------------------------------------------------------
// bin_start (exclusive scan of the nb + 1 counts) and the list of non-empty
// bricks, in two passes over tiles of SCAN_TILE counts: per-tile sums of the
// counts and of the non-empty flags, then per tile the sums of the tiles
// before it and a block scan of its own values.  A count and a flag travel
// as one 64-bit value (flag count high, particle count low: a particle
// total below 2^32 never carries into the high word).
__device__ __forceinline__ uint64_t block_exclusive_scan64(uint64_t v, uint64_t* total)
{
    __shared__ uint64_t wsum[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
        before += (w < wave) ? wsum[w] : 0u;
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// Count and non-empty flag of entry i of the nb + 1 counts (entry nb, the
// particles outside the bins, is no brick).
__device__ __forceinline__ uint64_t count_pair(const uint32_t* count, int64_t i, int64_t nb)
{
    if (i > nb) return 0u;
    const uint32_t c = count[i];
    return uint64_t(c) | (uint64_t(i < nb && c > 0u ? 1u : 0u) << 32);
}

__global__ __launch_bounds__(SCAN_THREADS) void k_tile_sums(const uint32_t* count, int64_t nb,
                                                            uint64_t* sums)
{
    const int64_t base = int64_t(blockIdx.x) * SCAN_TILE + int64_t(threadIdx.x) * SCAN_ITEMS;
    uint64_t s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) s += count_pair(count, base + q, nb);
    uint64_t total;
    (void)block_exclusive_scan64(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// In-place exclusive scan of the tile sums by one workgroup, for grids of
// more than SCAN_TILE tiles (fewer are summed by each tile directly).
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_sums(uint64_t* sums, int64_t n)
{
    uint64_t carry = 0;
    for (int64_t c0 = 0; c0 < n; c0 += SCAN_TILE) {
        const int64_t base = c0 + int64_t(threadIdx.x) * SCAN_ITEMS;
        uint64_t v[SCAN_ITEMS];
        uint64_t s = 0;
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) {
            v[q] = base + q < n ? sums[base + q] : 0u;
            s += v[q];
        }
        uint64_t total;
        uint64_t run = carry + block_exclusive_scan64(s, &total);
#pragma unroll
        for (int q = 0; q < SCAN_ITEMS; ++q) {
            if (base + q < n) sums[base + q] = run;
            run += v[q];
        }
        carry += total;
    }
}

// With items not null, non-empty brick t of the list also gets its walk
// item {t, brick, start, end}.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_compact(const uint32_t* count, int64_t nb,
                                                               const uint64_t* sums,
                                                               bool prefixed, uint32_t* bin_start,
                                                               uint32_t* active, BrickItem* items)
{
    // offset of this tile: the sum of the tiles before it
    uint64_t tile_off;
    if (prefixed) {
        tile_off = sums[blockIdx.x];
    } else {
        uint64_t off = 0;
        for (uint32_t u = threadIdx.x; u < blockIdx.x; u += SCAN_THREADS) off += sums[u];
        (void)block_exclusive_scan64(off, &tile_off);
    }
    const int64_t base = int64_t(blockIdx.x) * SCAN_TILE + int64_t(threadIdx.x) * SCAN_ITEMS;
    uint64_t v[SCAN_ITEMS];
    uint64_t s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        v[q] = count_pair(count, base + q, nb);
        s += v[q];
    }
    uint64_t total;
    uint64_t run = tile_off + block_exclusive_scan64(s, &total);
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        const int64_t i = base + q;
        if (i <= nb) {
            bin_start[i] = uint32_t(run);
            if (v[q] >> 32) {
                const uint32_t t = uint32_t(run >> 32);
                active[1 + t] = uint32_t(i);
                if (items)
                    items[t] = BrickItem{t, uint32_t(i), uint32_t(run),
                                         uint32_t(run) + uint32_t(v[q])};
            }
        }
        run += v[q];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_THREADS - 1)
        active[0] = uint32_t(run >> 32);
}


// ---------------------------------------------------------------------------
// Keys already non-decreasing (flag down): the stable sort is the identity.
// Otherwise every particle index goes to its brick segment of tmp, its key
// to the same slot of tmpk; the outsiders (key 0xffffffff) fill the segment
// after the last brick, in an order fixed by the brick pass.
__global__ __launch_bounds__(SORT_PT) void k_bin_scatter(int64_t np, const uint32_t* keys,
                                                         int32_t vshift, uint32_t bvol,
                                                         uint32_t nbricks, const uint32_t* start,
                                                         const uint32_t* unsorted,
                                                         uint32_t* cursor, uint32_t* tmp,
                                                         uint32_t* tmpk, uint32_t* perm)
{
    const int64_t p0 = int64_t(blockIdx.x) * (SORT_PT * SORT_ITEMS) + threadIdx.x;
    if (*unsorted == 0u) {
#pragma unroll
        for (int k = 0; k < SORT_ITEMS; ++k) {
            const int64_t p = p0 + k * SORT_PT;
            if (p < np) perm[p] = uint32_t(p);
        }
        return;
    }
    uint32_t key[SORT_ITEMS];
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
        const int64_t p = p0 + k * SORT_PT;
        key[k] = p < np ? keys[p] : 0xffffffffu;
    }
    AggTicket tk[SORT_ITEMS];
    uint32_t at[SORT_ITEMS];
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
        const bool live = p0 + k * SORT_PT < np;
        const uint32_t b = key[k] == 0xffffffffu ? nbricks
                           : vshift >= 0         ? key[k] >> vshift
                                                 : key[k] / bvol;
        at[k] = live ? start[b] : 0u;
        tk[k] = wave_agg_issue(cursor, b, live);
    }
#pragma unroll
    for (int k = 0; k < SORT_ITEMS; ++k) {
        const int64_t p = p0 + k * SORT_PT;
        const uint32_t r = wave_agg_resolve(tk[k]);
        if (p >= np) continue;
        tmp[at[k] + r] = uint32_t(p);
        tmpk[at[k] + r] = key[k];
    }
}

// Sort one brick segment held in `items` (LDS or global scratch): count the
// bins, scan, place each index in its bin (unordered), then give each index
// its rank among the indices of its bin, which is its stable position.  The
// segment's indices are tmp[beg, beg + n), their keys tmpk[beg, beg + n).
__device__ void brick_segment_sort(uint32_t b, uint32_t beg, uint32_t n, uint32_t bvol,
                                   const uint32_t* tmp, const uint32_t* tmpk,
                                   uint32_t* items, uint32_t* cnt, uint32_t* cur,
                                   uint32_t* perm, uint32_t* keys_out)
{
    const uint32_t kb = b * bvol;
    for (uint32_t l = threadIdx.x; l < bvol; l += blockDim.x) cnt[l] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) atomicAdd(&cnt[tmpk[beg + i] - kb], 1u);
    __syncthreads();
    // exclusive scan of cnt into cur (one wave, 64 bins at a time)
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        uint32_t carry = 0;
        for (uint32_t c0 = 0; c0 < bvol; c0 += 64) {
            const uint32_t v = (c0 + lane < bvol) ? cnt[c0 + lane] : 0u;
            uint32_t inc = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(inc, o);
                if (lane >= o) inc += y;
            }
            if (c0 + lane < bvol) cur[c0 + lane] = carry + inc - v;
            carry += __shfl(inc, 63);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t pos = atomicAdd(&cur[tmpk[beg + i] - kb], 1u);
        items[pos] = tmp[beg + i];
    }
    __syncthreads();
    // cur[l] is now the end of bin l; item i lies in the first bin whose end
    // exceeds i.  Indices are distinct, so the ranks inside a bin are a
    // permutation of its slots.
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t v = items[i];
        uint32_t lo = 0, hi = bvol - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cur[mid] > i) hi = mid;
            else lo = mid + 1;
        }
        const uint32_t l = lo;
        const uint32_t e = cur[l];
        uint32_t r = e - cnt[l];
        for (uint32_t j = e - cnt[l]; j < e; ++j) r += items[j] < v ? 1u : 0u;
        perm[beg + r] = v;
        keys_out[beg + r] = kb + l;
    }
}

// brick_segment_sort for segments of at most SORT_REG * blockDim.x
// indices: each index and its key are loaded once (coalesced) into
// registers; the bin count's LDS atomic returns the index's arrival rank,
// which places it in its bin with no second pass over memory, and the bin
// of every placed index is kept in LDS (ibin) for the rank step.
constexpr int SORT_REG = 4;
static_assert(SORT_REG * SORT_BT <= SORT_LDS_CAP, "register-path segments fit the LDS arrays");

__device__ void brick_segment_sort_reg(uint32_t b, uint32_t beg, uint32_t n, uint32_t bvol,
                                       const uint32_t* tmp, const uint32_t* tmpk,
                                       uint32_t* items, uint32_t* ibin, uint32_t* cnt,
                                       uint32_t* cur, uint32_t* perm, uint32_t* keys_out)
{
    const uint32_t kb = b * bvol;
    uint32_t pv[SORT_REG], lb[SORT_REG], rk[SORT_REG];
#pragma unroll
    for (int r = 0; r < SORT_REG; ++r) {
        const uint32_t i = threadIdx.x + uint32_t(r) * blockDim.x;
        pv[r] = i < n ? tmp[beg + i] : 0u;
        lb[r] = i < n ? tmpk[beg + i] - kb : 0u;
    }
    for (uint32_t l = threadIdx.x; l < bvol; l += blockDim.x) cnt[l] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_REG; ++r) {
        const uint32_t i = threadIdx.x + uint32_t(r) * blockDim.x;
        if (i < n) rk[r] = atomicAdd(&cnt[lb[r]], 1u);
    }
    __syncthreads();
    // exclusive scan of cnt into cur (one wave, 64 bins at a time)
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        uint32_t carry = 0;
        for (uint32_t c0 = 0; c0 < bvol; c0 += 64) {
            const uint32_t v = (c0 + lane < bvol) ? cnt[c0 + lane] : 0u;
            uint32_t inc = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(inc, o);
                if (lane >= o) inc += y;
            }
            if (c0 + lane < bvol) cur[c0 + lane] = carry + inc - v;
            carry += __shfl(inc, 63);
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_REG; ++r) {
        const uint32_t i = threadIdx.x + uint32_t(r) * blockDim.x;
        if (i < n) {
            const uint32_t pos = cur[lb[r]] + rk[r];
            items[pos] = pv[r];
            ibin[pos] = lb[r];
        }
    }
    __syncthreads();
    // Indices are distinct, so the ranks inside a bin are a permutation of
    // its slots.
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t v = items[i];
        const uint32_t l = ibin[i];
        const uint32_t s0 = cur[l], e = s0 + cnt[l];
        uint32_t r = s0;
        for (uint32_t j = s0; j < e; ++j) r += items[j] < v ? 1u : 0u;
        perm[beg + r] = v;
        keys_out[beg + r] = kb + l;
    }
}

// Workgroups walk the non-empty bricks with a grid stride; the segment of
// particles outside the bins, [start[nbricks], np), is copied as it was
// scattered (key 0xffffffff).  Nothing to do when the keys arrived sorted.
// Dynamic LDS: cnt[bvol] | cur[bvol] | items[SORT_LDS_CAP] |
// ibin[SORT_LDS_CAP]; bricks of more than SORT_LDS_CAP indices use `big`.
__global__ __launch_bounds__(SORT_BT) void k_brick_sort(int64_t np, uint32_t nbricks,
                                                        const uint32_t* start,
                                                        const uint32_t* active,
                                                        const uint32_t* unsorted,
                                                        const uint32_t* tmp,
                                                        const uint32_t* tmpk, uint32_t bvol,
                                                        uint32_t* big, uint32_t* perm,
                                                        uint32_t* keys_out)
{
    if (*unsorted == 0u) return;
    extern __shared__ uint32_t sort_lds[];
    uint32_t* cnt = sort_lds;
    uint32_t* cur = sort_lds + bvol;
    uint32_t* lds_items = sort_lds + 2 * bvol;
    uint32_t* ibin = lds_items + SORT_LDS_CAP;
    for (int64_t s = int64_t(start[nbricks]) + int64_t(blockIdx.x) * SORT_BT + threadIdx.x;
         s < np; s += int64_t(gridDim.x) * SORT_BT) {
        perm[s] = tmp[s];
        keys_out[s] = 0xffffffffu;
    }
    const uint32_t nact = active[0];
    for (uint32_t t = blockIdx.x; t < nact; t += gridDim.x) {
        const uint32_t b = active[1 + t];
        const uint32_t beg = start[b], n = start[b + 1] - beg;
        if (n <= uint32_t(SORT_REG * SORT_BT)) {
            brick_segment_sort_reg(b, beg, n, bvol, tmp, tmpk, lds_items, ibin, cnt, cur, perm,
                                   keys_out);
        } else {
            uint32_t* items = n <= SORT_LDS_CAP ? lds_items : big + beg;
            brick_segment_sort(b, beg, n, bvol, tmp, tmpk, items, cnt, cur, perm, keys_out);
        }
        __syncthreads();
    }
}

}  // namespace pic

using namespace pic;

extern "C" int64_t pic_bin_count(const pic_binning* g)
{
    if (!g) return PIC_ERR_ARG;
    int64_t nb = 1, bvol = 1;
    for (int d = 0; d < 3; ++d) {
        if (g->ncell[d] <= 0 || g->brick[d] <= 0) return PIC_ERR_ARG;
        nb *= (g->ncell[d] + g->brick[d] - 1) / g->brick[d];
        bvol *= g->brick[d];
    }
    if (dual_mask(g->dual) < 0) return PIC_ERR_ARG;
    if (bvol > SORT_MAX_BVOL || nb * bvol >= int64_t(UINT32_MAX) || nb >= int64_t(INT32_MAX))
        return PIC_ERR_CAPACITY;
    return nb;
}

namespace pic {

KeyGeom make_key_geom(const pic_mesh& mesh, const pic_binning& b)
{
    KeyGeom g;
    g.bvol = 1;
    for (int d = 0; d < 3; ++d) {
        g.low[d] = mesh.low[d];
        g.cell_lo[d] = b.cell_lo[d];
        g.ncell[d] = b.ncell[d];
        g.brick[d] = b.brick[d];
        g.nb[d] = (b.ncell[d] + b.brick[d] - 1) / b.brick[d];
        g.bvol *= b.brick[d];
    }
    g.rdx = 1.0 / mesh.h;
    g.dual = dual_mask(b.dual);
    g.vshift = 0;
    for (int d = 0; d < 3; ++d) {
        const int e = b.brick[d];
        int sh = 0;
        while ((1 << sh) < e) ++sh;
        g.bshift[d] = (1 << sh) == e ? sh : -1;
        if (g.bshift[d] < 0) g.vshift = -1;
        else if (g.vshift >= 0) g.vshift += sh;
    }
    return g;
}

}  // namespace pic

namespace pic {

static int check_sort_args(pic_context* ctx, int dtype, const pic_mesh* mesh, int64_t np,
                           const pic_pfield* x, const pic_binning* geom, const uint32_t* keys,
                           const uint32_t* perm, const uint32_t* bin_start, int64_t* nb)
{
    if (!ctx || !mesh || !(mesh->h > 0.0) || (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 ||
        np >= int64_t(UINT32_MAX) || !geom || !bin_start)
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !keys || !perm)) return PIC_ERR_ARG;
    *nb = pic_bin_count(geom);
    return *nb < 0 ? int(*nb) : PIC_OK;
}

SortLayout sort_layout(int64_t nb, int64_t np, bool caller_active)
{
    SortLayout l;
    l.ntile = (nb + 1 + SCAN_TILE - 1) / SCAN_TILE;
    l.nzero = 2 * (nb + 1) + 2;
    l.head = l.nzero + 2 * l.ntile + (caller_active ? 0 : nb + 1);
    l.bytes = size_t(l.head + 3 * np) * 4;
    return l;
}

void sort_clear(pic_context* ctx, const SortLayout& L, uint32_t* sc)
{
    const unsigned fb = (unsigned)std::min<int64_t>((L.nzero + 255) / 256, 4096);
    k_fill_u32<<<fb, 256, 0, ctx->stream>>>(sc, L.nzero, 0u);
}

uint32_t* sort_active(const SortLayout& L, uint32_t* sc)
{
    return reinterpret_cast<uint32_t*>(reinterpret_cast<uint64_t*>(sc + L.nzero) + L.ntile);
}

int sort_finish(pic_context* ctx, int64_t np, int64_t nb, uint32_t bvol, int32_t vshift,
                const SortLayout& L, uint32_t* sc, uint32_t* keys, uint32_t* perm,
                uint32_t* bin_start, uint32_t* active, BrickItem* items)
{
    uint32_t* count = sc;
    uint32_t* cursor = count + nb + 1;
    uint32_t* unsorted = cursor + nb + 1;
    uint64_t* tile_sums = reinterpret_cast<uint64_t*>(sc + L.nzero);
    uint32_t* act = active ? active : sort_active(L, sc);
    uint32_t* tmp = sc + L.head;
    uint32_t* tmpk = tmp + np;
    uint32_t* big = tmpk + np;
    hipStream_t st = ctx->stream;
    k_tile_sums<<<(unsigned)L.ntile, SCAN_THREADS, 0, st>>>(count, nb, tile_sums);
    const bool prefixed = L.ntile > SCAN_TILE;
    if (prefixed) k_scan_sums<<<1, SCAN_THREADS, 0, st>>>(tile_sums, L.ntile);
    k_scan_compact<<<(unsigned)L.ntile, SCAN_THREADS, 0, st>>>(count, nb, tile_sums, prefixed,
                                                               bin_start, act, items);
    if (np > 0) {
        k_bin_scatter<<<blocks_for(np, SORT_PT * SORT_ITEMS), SORT_PT, 0, st>>>(
            np, keys, vshift, bvol, uint32_t(nb), bin_start, unsorted, cursor, tmp, tmpk, perm);
        const size_t lds = size_t(2 * bvol + 2 * SORT_LDS_CAP) * 4;
        const unsigned nbk = persistent_blocks(ctx, k_brick_sort, SORT_BT, lds, nb + 1);
        k_brick_sort<<<nbk, SORT_BT, lds, st>>>(np, uint32_t(nb), bin_start, act, unsorted, tmp,
                                                tmpk, bvol, big, perm, keys);
    }
    return launch_status();
}

// The sort's passes on the context's stream with scratch `sc` laid out by
// sort_layout (device memory of at least L.bytes), the clearing job z run
// beside the key pass and, when items is not null, the walk items listed
// by the scan.
int bin_sort_with(pic_context* ctx, int dtype, const pic_mesh& mesh, int64_t np,
                  const pic_pfield* x, const pic_binning& geom, int64_t nb, const SortLayout& L,
                  uint32_t* sc, uint32_t* keys, uint32_t* perm, uint32_t* bin_start,
                  uint32_t* active, const ZeroJob& z, BrickItem* items)
{
    const KeyGeom g = make_key_geom(mesh, geom);
    sort_clear(ctx, L, sc);
    uint32_t* count = sort_count(sc);
    uint32_t* unsorted = sort_unsorted(sc, nb);
    if (z.blocks > 0) {
        const unsigned b = key_blocks(np) + z.blocks;
        if (dtype == PIC_F64)
            k_cell_keys_clear<<<b, SORT_PT, 0, ctx->stream>>>(
                np, MeshKeys<double>{make_pfield<double>(*x), g}, uint32_t(nb), keys, count,
                unsorted, ctx->d_err, z);
        else
            k_cell_keys_clear<<<b, SORT_PT, 0, ctx->stream>>>(
                np, MeshKeys<float>{make_pfield<float>(*x), g}, uint32_t(nb), keys, count,
                unsorted, ctx->d_err, z);
    } else if (np > 0) {
        const unsigned b = key_blocks(np);
        if (dtype == PIC_F64)
            k_cell_keys<<<b, SORT_PT, 0, ctx->stream>>>(
                np, MeshKeys<double>{make_pfield<double>(*x), g}, uint32_t(nb), keys, count,
                unsorted, ctx->d_err);
        else
            k_cell_keys<<<b, SORT_PT, 0, ctx->stream>>>(
                np, MeshKeys<float>{make_pfield<float>(*x), g}, uint32_t(nb), keys, count,
                unsorted, ctx->d_err);
    }
    return sort_finish(ctx, np, nb, g.bvol, g.vshift, L, sc, keys, perm, bin_start, active,
                       items);
}

static int bin_sort_run(pic_context* ctx, int dtype, const pic_mesh& mesh, int64_t np,
                        const pic_pfield* x, const pic_binning& geom, int64_t nb,
                        const SortLayout& L, uint32_t* sc, uint32_t* keys, uint32_t* perm,
                        uint32_t* bin_start, uint32_t* active)
{
    return bin_sort_with(ctx, dtype, mesh, np, x, geom, nb, L, sc, keys, perm, bin_start, active,
                         ZeroJob{}, nullptr);
}

}  // namespace pic

extern "C" int pic_bin_sort(pic_context* ctx, int dtype, const pic_mesh* mesh, int64_t np,
                            const pic_pfield* x, const pic_binning* geom, uint32_t* keys,
                            uint32_t* perm, uint32_t* bin_start, uint32_t* active)
{
    Region region("picasso_amd::bin_sort");
    int64_t nb = 0;
    if (const int rc = check_sort_args(ctx, dtype, mesh, np, x, geom, keys, perm, bin_start, &nb))
        return rc;
    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;
    const SortLayout L = sort_layout(nb, np, active != nullptr);
    uint32_t* sc = static_cast<uint32_t*>(pic::scratch(ctx, L.bytes));
    if (!sc) return PIC_ERR_ALLOC;
    TimedLaunch tl(ctx, "bin_sort");
    return bin_sort_run(ctx, dtype, *mesh, np, x, *geom, nb, L, sc, keys, perm, bin_start, active);
}

// ---------------------------------------------------------------------------
// Particle fields gathered into bin order: dst(s, c) = src(perm[s], c).
namespace pic {

template <typename T>
__global__ __launch_bounds__(256) void k_permute(int64_t np, const uint32_t* perm, PField<T> src,
                                                 PField<T> dst, int ncomp)
{
    for (int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; s < np;
         s += int64_t(gridDim.x) * blockDim.x) {
        const int64_t sb = src.base(int64_t(perm[s])), db = dst.base(s);
        for (int c = 0; c < ncomp; ++c) dst.at(db, c) = src.at(sb, c);
    }
}

}  // namespace pic

extern "C" int pic_permute(pic_context* ctx, int dtype, int64_t np, const uint32_t* perm,
                           int nfields, const pic_pfield* src, const pic_pfield* dst,
                           const int* ncomp)
{
    Region region("picasso_amd::permute");
    if (!ctx || (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 || nfields < 0 ||
        (nfields > 0 && (!src || !dst || !ncomp)))
        return PIC_ERR_ARG;
    if (np == 0 || nfields == 0) return PIC_OK;
    if (!perm) return PIC_ERR_ARG;
    for (int f = 0; f < nfields; ++f)
        if (!valid_pfield(&src[f]) || !valid_pfield(&dst[f]) || ncomp[f] < 1) return PIC_ERR_ARG;
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    TimedLaunch tl(ctx, "permute");
    const unsigned b = (unsigned)std::min<int64_t>(blocks_for(np, 256), 65536);
    for (int f = 0; f < nfields; ++f) {
        if (dtype == PIC_F64)
            k_permute<double><<<b, 256, 0, ctx->stream>>>(np, perm, make_pfield<double>(src[f]),
                                                          make_pfield<double>(dst[f]), ncomp[f]);
        else
            k_permute<float><<<b, 256, 0, ctx->stream>>>(np, perm, make_pfield<float>(src[f]),
                                                         make_pfield<float>(dst[f]), ncomp[f]);
    }
    return launch_status();
}

// ---------------------------------------------------------------------------
// Stable selection: the list of the particles a mask keeps, in increasing
// order, by the sort's tile-sum / scan-compact passes over 0/1 counts.
namespace pic {

__global__ void k_keep_counts(int64_t np, const uint8_t* keep, uint32_t* count)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i <= np;
         i += int64_t(gridDim.x) * blockDim.x)
        count[i] = i < np && keep[i] ? 1u : 0u;
}

}  // namespace pic

extern "C" int pic_select(pic_context* ctx, int64_t np, const uint8_t* keep, uint32_t* kept)
{
    Region region("picasso_amd::select");
    if (!ctx || np < 0 || np >= int64_t(UINT32_MAX) || !kept || (np > 0 && !keep))
        return PIC_ERR_ARG;
    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;
    // scratch: count[np+1] | scan[np+1] | tile sums (64-bit)
    const int64_t ntile = (np + 1 + SCAN_TILE - 1) / SCAN_TILE;
    const int64_t head = (2 * (np + 1) + 1) & ~int64_t(1);  // tile sums 8-byte aligned
    uint32_t* sc = static_cast<uint32_t*>(scratch(ctx, size_t(head + 2 * ntile) * 4));
    if (!sc) return PIC_ERR_ALLOC;
    uint32_t* count = sc;
    uint32_t* scan = sc + np + 1;
    uint64_t* tile_sums = reinterpret_cast<uint64_t*>(sc + head);
    hipStream_t st = ctx->stream;
    TimedLaunch tl(ctx, "select");
    const unsigned b = (unsigned)std::min<int64_t>(blocks_for(np + 1, 256), 65536);
    k_keep_counts<<<b, 256, 0, st>>>(np, keep, count);
    k_tile_sums<<<(unsigned)ntile, SCAN_THREADS, 0, st>>>(count, np, tile_sums);
    const bool prefixed = ntile > SCAN_TILE;
    if (prefixed) k_scan_sums<<<1, SCAN_THREADS, 0, st>>>(tile_sums, ntile);
    k_scan_compact<<<(unsigned)ntile, SCAN_THREADS, 0, st>>>(count, np, tile_sums, prefixed, scan,
                                                             kept, nullptr);
    return launch_status();
}

// ---------------------------------------------------------------------------
// Allocation-free bin sort: the same passes with their scratch in a
// file-scope device array of PIC_NOALLOC_SORT_BYTES, which every GPU the
// library runs on holds until the process ends.  In host code the array's
// name is not a device address: the address is resolved per device with
// hipGetSymbolAddress and only that pointer is handed to kernels.
static __device__ __align__(256) uint32_t scratch[PIC_NOALLOC_SORT_BYTES / 4];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
void* noalloc_ptr[NOALLOC_MAX_DEVICES];  // device address of `scratch`, per device
}  // namespace

namespace pic {

int noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return PIC_ERR_ARG;
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_ptr[device]) return PIC_OK;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) return PIC_ERR_HIP;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)) != hipSuccess || !p) return PIC_ERR_HIP;
    noalloc_ptr[device] = p;
    return PIC_OK;
}

}  // namespace pic

extern "C" int pic_bin_sort_noalloc(pic_context* ctx, int dtype, const pic_mesh* mesh, int64_t np,
                                    const pic_pfield* x, const pic_binning* geom, uint32_t* keys,
                                    uint32_t* perm, uint32_t* bin_start, uint32_t* active)
{
    Region region("picasso_amd::bin_sort_noalloc");
    int64_t nb = 0;
    if (const int rc = check_sort_args(ctx, dtype, mesh, np, x, geom, keys, perm, bin_start, &nb))
        return rc;
    const SortLayout L = sort_layout(nb, np, active != nullptr);
    if (L.bytes > sizeof(::scratch)) return PIC_ERR_CAPACITY;
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES) return PIC_ERR_ARG;
    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;
    if (const int rc = noalloc_setup(ctx->device)) return rc;
    void* p = nullptr;
    {
        std::lock_guard<std::mutex> lk(noalloc_mu);
        p = noalloc_ptr[ctx->device];
    }
    return bin_sort_run(ctx, dtype, *mesh, np, x, *geom, nb, L, static_cast<uint32_t*>(p), keys,
                        perm, bin_start, active);
}
