// Cell-binning particle sort (no reference counterpart: Picasso never sorts
// particles).  Stable counting sort by bin key:
//   key = brick_linear * brick_volume + bin_in_brick_linear (both i-major),
// so a brick's particles are contiguous and, inside it, ordered by bin and
// then by input index.
//
// Passes
//  (1) keys in input order + per-brick counts (wave-aggregated atomics);
//  (2) exclusive scan of the counts into bin_start;
//  (3) scatter of particle indices into their brick segments (wave-aggregated
//      cursors: order inside a segment depends on scheduling);
//  (4) one workgroup per brick: LDS counting sort by bin, then each bin's
//      indices sorted ascending, which restores the stable order bit-exactly
//      whatever order pass 3 produced.  Writes perm and the sorted keys.
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

// Wave-aggregated atomicAdd of 1 per active lane on base[key]: lanes with
// equal keys share one atomic.  Returns each lane's slot (old value + rank of
// the lane among equal-key lanes, in lane order).
__device__ __forceinline__ uint32_t wave_agg_inc(uint32_t* base, uint32_t key, bool active)
{
    const int lane = __lane_id();
    uint64_t pending = __ballot(active);
    uint32_t slot = 0;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lkey = __shfl(key, leader);
        const uint64_t same = __ballot(active && key == lkey) & pending;
        uint32_t b = 0;
        if (lane == leader) b = atomicAdd(base + lkey, (uint32_t)__popcll(same));
        b = __shfl(b, leader);
        if ((same >> lane) & 1ull) slot = b + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        pending &= ~same;
    }
    return slot;
}

// Particles outside the bins are counted in count[nbricks] (they are placed
// after every brick, so perm stays a permutation even then).
template <typename T>
__global__ __launch_bounds__(256) void k_cell_keys(int64_t np, PField<T> x, KeyGeom g,
                                                   uint32_t nbricks, uint32_t* keys,
                                                   uint32_t* count, int* err)
{
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool live = p < np;
    uint32_t key = 0xffffffffu, b = 0;
    if (live) {
        key = cell_key(g, x, p, &b);
        keys[p] = key;
        if (key == 0xffffffffu) {
            atomicOr(err, ERRBIT_CELL);
            b = nbricks;
        }
    }
    (void)wave_agg_inc(count, b, live);
}

// ---------------------------------------------------------------------------
// Device-wide exclusive scan of uint32.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t wsum[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SCAN_THREADS / 64; ++w) {
        before += (w < wave) ? wsum[w] : 0u;
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_tile(const uint32_t* in, uint32_t* out,
                                                            int64_t n, uint32_t* sums)
{
    const int64_t base = int64_t(blockIdx.x) * SCAN_TILE + int64_t(threadIdx.x) * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        v[q] = (base + q < n) ? in[base + q] : 0u;
        s += v[q];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, &total);
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == 0 && sums) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_add(uint32_t* out, int64_t n,
                                                           const uint32_t* offs)
{
    const int64_t base = int64_t(blockIdx.x) * SCAN_TILE;
    const uint32_t add = offs[blockIdx.x];
    for (int q = threadIdx.x; q < SCAN_TILE; q += SCAN_THREADS)
        if (base + q < n) out[base + q] += add;
}

// Words of temporary storage exclusive_scan needs for n values.
int64_t scan_tmp_words(int64_t n)
{
    int64_t w = 0;
    while (n > SCAN_TILE) {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        w += n + 1;
    }
    return w + 1;
}

void exclusive_scan(hipStream_t st, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp)
{
    const int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nt <= 1) {
        k_scan_tile<<<1, SCAN_THREADS, 0, st>>>(in, out, n, nullptr);
        return;
    }
    uint32_t* sums = tmp;
    k_scan_tile<<<(unsigned)nt, SCAN_THREADS, 0, st>>>(in, out, n, sums);
    exclusive_scan(st, sums, sums, nt, tmp + nt + 1);
    k_scan_add<<<(unsigned)nt, SCAN_THREADS, 0, st>>>(out, n, sums);
}

// ---------------------------------------------------------------------------
// Particles outside the bins go straight to their final slots after the
// last brick (key 0xffffffff, order unspecified: the call reports an error).
__global__ __launch_bounds__(256) void k_bin_scatter(int64_t np, const uint32_t* keys,
                                                     uint32_t bvol, uint32_t nbricks,
                                                     const uint32_t* start, uint32_t* cursor,
                                                     uint32_t* tmp, uint32_t* perm,
                                                     uint32_t* keys_out)
{
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool live = p < np;
    const uint32_t key = live ? keys[p] : 0xffffffffu;
    const bool ok = key != 0xffffffffu;
    const uint32_t b = ok ? key / bvol : nbricks;
    const uint32_t r = wave_agg_inc(cursor, b, live);
    if (live && ok) tmp[start[b] + r] = uint32_t(p);
    if (live && !ok) {
        perm[start[nbricks] + r] = uint32_t(p);
        keys_out[start[nbricks] + r] = 0xffffffffu;
    }
}

// Sort one brick segment held in `items` (LDS or global scratch): count the
// bins, scan, place each index in its bin (unordered), then give each index
// its rank among the indices of its bin, which is its stable position.
__device__ void brick_segment_sort(uint32_t b, uint32_t beg, uint32_t n, uint32_t bvol,
                                   const uint32_t* tmp, const uint32_t* keys_in,
                                   uint32_t* items, uint32_t* cnt, uint32_t* cur,
                                   uint32_t* perm, uint32_t* keys_out)
{
    const uint32_t kb = b * bvol;
    for (uint32_t l = threadIdx.x; l < bvol; l += blockDim.x) cnt[l] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        atomicAdd(&cnt[keys_in[tmp[beg + i]] - kb], 1u);
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
        const uint32_t p = tmp[beg + i];
        const uint32_t pos = atomicAdd(&cur[keys_in[p] - kb], 1u);
        items[pos] = p;
    }
    __syncthreads();
    // cur[l] is now the end of bin l.  Indices are distinct, so the ranks
    // inside a bin are a permutation of its slots.
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint32_t v = items[i];
        const uint32_t l = keys_in[v] - kb;
        const uint32_t e = cur[l];
        uint32_t r = e - cnt[l];
        for (uint32_t j = e - cnt[l]; j < e; ++j) r += items[j] < v ? 1u : 0u;
        perm[beg + r] = v;
        keys_out[beg + r] = kb + l;
    }
}

// List of non-empty bricks: flags -> exclusive scan -> ids in increasing order.
__global__ __launch_bounds__(256) void k_nonempty_flags(const uint32_t* count, int64_t nb,
                                                        uint32_t* flags)
{
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b <= nb) flags[b] = (b < nb && count[b] > 0u) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_compact_active(const uint32_t* flags, const uint32_t* pos,
                                                        int64_t nb, uint32_t* active)
{
    const int64_t b = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b < nb && flags[b]) active[1 + pos[b]] = uint32_t(b);
    if (b == nb) active[0] = pos[nb];
}

// Workgroups walk the non-empty bricks with a grid stride.  Dynamic LDS:
// cnt[bvol] | cur[bvol] | items[SORT_LDS_CAP]; larger bricks use `big`.
__global__ __launch_bounds__(SORT_BT) void k_brick_sort(const uint32_t* start,
                                                        const uint32_t* active,
                                                        const uint32_t* tmp,
                                                        const uint32_t* keys_in, uint32_t bvol,
                                                        uint32_t* big, uint32_t* perm,
                                                        uint32_t* keys_out)
{
    extern __shared__ uint32_t sort_lds[];
    uint32_t* cnt = sort_lds;
    uint32_t* cur = sort_lds + bvol;
    uint32_t* lds_items = sort_lds + 2 * bvol;
    const uint32_t nact = active[0];
    for (uint32_t t = blockIdx.x; t < nact; t += gridDim.x) {
        const uint32_t b = active[1 + t];
        const uint32_t beg = start[b], n = start[b + 1] - beg;
        uint32_t* items = n <= SORT_LDS_CAP ? lds_items : big + beg;
        brick_segment_sort(b, beg, n, bvol, tmp, keys_in, items, cnt, cur, perm, keys_out);
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
    if (g->dual != 0 && g->dual != 1) return PIC_ERR_ARG;
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
    g.dual = b.dual;
    return g;
}

}  // namespace pic

extern "C" int pic_bin_sort(pic_context* ctx, int dtype, const pic_mesh* mesh, int64_t np,
                            const pic_pfield* x, const pic_binning* geom, uint32_t* keys,
                            uint32_t* perm, uint32_t* bin_start, uint32_t* active)
{
    if (!ctx || !mesh || !(mesh->h > 0.0) || (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 ||
        np >= int64_t(UINT32_MAX) || !geom || !bin_start)
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !keys || !perm)) return PIC_ERR_ARG;
    const int64_t nb = pic_bin_count(geom);
    if (nb < 0) return (int)nb;
    const KeyGeom g = make_key_geom(*mesh, *geom);
    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;

    // scratch: count[nb+1] | cursor[nb+1] | scan tmp | act[nb+1] | keys_in[np] | tmp[w] | big[w]
    // with w = max(np, nb+1) (tmp and big also hold the non-empty flags and
    // their scan before the scatter).  act holds the active-brick list when
    // the caller does not ask for it.
    const int64_t scan_words = scan_tmp_words(nb + 1);
    const int64_t head = (nb + 1) + (nb + 1) + scan_words + (active ? 0 : nb + 1);
    const int64_t w = std::max<int64_t>(np, nb + 1);
    const size_t bytes = size_t(head + np + 2 * w) * 4;
    uint32_t* sc = static_cast<uint32_t*>(scratch(ctx, bytes));
    if (!sc) return PIC_ERR_ALLOC;
    uint32_t* count = sc;
    uint32_t* cursor = count + nb + 1;
    uint32_t* scan_tmp = cursor + nb + 1;
    uint32_t* act = active ? active : scan_tmp + scan_words;
    uint32_t* keys_in = sc + head;
    uint32_t* tmp = keys_in + np;
    uint32_t* big = tmp + w;
    hipStream_t st = ctx->stream;

    TimedLaunch tl(ctx, "bin_sort");
    const unsigned fb = (unsigned)std::min<int64_t>((2 * nb + 2 + 255) / 256, 4096);
    k_fill_u32<<<fb, 256, 0, st>>>(count, 2 * nb + 2, 0u);
    if (np > 0) {
        const unsigned b = blocks_for(np, 256);
        if (dtype == PIC_F64)
            k_cell_keys<double><<<b, 256, 0, st>>>(np, make_pfield<double>(*x), g, uint32_t(nb),
                                                    keys_in, count,
                                                    ctx->d_err);
        else
            k_cell_keys<float><<<b, 256, 0, st>>>(np, make_pfield<float>(*x), g, uint32_t(nb),
                                                   keys_in, count,
                                                   ctx->d_err);
    }
    exclusive_scan(st, count, bin_start, nb + 1, scan_tmp);
    {
        // non-empty brick list: flags and their scan live in tmp and big
        // before the scatter uses them.
        uint32_t* flags = tmp;
        uint32_t* pos = big;
        const unsigned fb2 = blocks_for(nb + 1, 256);
        k_nonempty_flags<<<fb2, 256, 0, st>>>(count, nb, flags);
        exclusive_scan(st, flags, pos, nb + 1, scan_tmp);
        k_compact_active<<<fb2, 256, 0, st>>>(flags, pos, nb, act);
    }
    if (np > 0) {
        k_bin_scatter<<<blocks_for(np, 256), 256, 0, st>>>(np, keys_in, g.bvol, uint32_t(nb),
                                                            bin_start, cursor, tmp, perm, keys);
        const size_t lds = size_t(2 * g.bvol + SORT_LDS_CAP) * 4;
        const unsigned nbk = persistent_blocks(ctx, k_brick_sort, SORT_BT, lds, nb);
        k_brick_sort<<<nbk, SORT_BT, lds, st>>>(bin_start, act, tmp, keys_in, g.bvol, big, perm,
                                                keys);
    }
    return launch_status();
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
