// Cell-binning particle sort (no reference counterpart: Picasso never sorts
// particles).  Stable counting sort by brick-major cell key:
//   key = brick_linear * brick_volume + cell_in_brick_linear (both i-major),
// so a brick's particles are contiguous and, inside it, ordered by cell and
// then by input index.
//
// Passes: (1) keys + per-brick counts (atomics), (2) exclusive scan of the
// counts into bin_start, (3) scatter of (cell_in_brick, index) pairs into the
// brick segments (atomic cursors, so unordered inside a segment), (4) one
// workgroup per brick sorts its segment in LDS by (cell_in_brick, index),
// which restores stability bit-exactly whatever order pass 3 produced.
#include "pic_common.hpp"
#include "pic_internal.hpp"

namespace pic {

constexpr int SORT_LDS_CAP = 8192;  // particles per brick sorted in LDS
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v)
{
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

struct KeyGeom {
    double low[3];
    double rdx;
    int32_t cell_lo[3];
    int32_t ncell[3];
    int32_t brick[3];
    int32_t nb[3];
    uint32_t bvol;
};

template <typename T>
__global__ __launch_bounds__(256) void k_cell_keys(int64_t np, PField<T> x, KeyGeom g,
                                                   uint32_t* keys, uint32_t* count, int* err)
{
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int64_t xb = x.base(p);
    int32_t rel[3];
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double t = ((double)x.at(xb, d) - g.low[d]) * g.rdx;
        const double c = floor(t) - (double)g.cell_lo[d];
        ok = ok && (c >= 0.0) && (c < (double)g.ncell[d]);
        rel[d] = ok ? (int32_t)c : 0;
    }
    if (!ok) {
        keys[p] = 0xffffffffu;
        atomicOr(err, ERRBIT_CELL);
        return;
    }
    const uint32_t bi = rel[0] / g.brick[0], bj = rel[1] / g.brick[1], bk = rel[2] / g.brick[2];
    const uint32_t li = rel[0] % g.brick[0], lj = rel[1] % g.brick[1], lk = rel[2] % g.brick[2];
    const uint32_t b = (bi * g.nb[1] + bj) * g.nb[2] + bk;
    const uint32_t l = (li * g.brick[1] + lj) * g.brick[2] + lk;
    keys[p] = b * g.bvol + l;
    atomicAdd(count + b, 1u);
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

// Exclusive scan of n values (in and out may alias); tmp needs
// scan_tmp_words(n) words.
static int64_t scan_tmp_words(int64_t n)
{
    int64_t w = 0;
    while (n > SCAN_TILE) {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        w += n + 1;
    }
    return w + 1;
}

static void exclusive_scan(hipStream_t st, const uint32_t* in, uint32_t* out, int64_t n,
                           uint32_t* tmp)
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
__global__ __launch_bounds__(256) void k_bin_scatter(int64_t np, const uint32_t* keys,
                                                     uint32_t bvol, const uint32_t* start,
                                                     uint32_t* cursor, uint64_t* tmp)
{
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const uint32_t key = keys[p];
    if (key == 0xffffffffu) return;
    const uint32_t b = key / bvol;
    const uint32_t slot = start[b] + atomicAdd(cursor + b, 1u);
    tmp[slot] = (uint64_t(key - b * bvol) << 32) | uint64_t(p);
}

// One workgroup per brick: sort its segment by (cell_in_brick, index).
__global__ __launch_bounds__(256) void k_brick_sort(const uint32_t* start, const uint64_t* tmp,
                                                    uint32_t* perm)
{
    extern __shared__ uint64_t s[];
    const uint32_t b = blockIdx.x;
    const uint32_t beg = start[b], n = start[b + 1] - beg;
    if (n == 0) return;
    if (n <= SORT_LDS_CAP) {
        uint32_t N = 1;
        while (N < n) N <<= 1;
        for (uint32_t i = threadIdx.x; i < N; i += blockDim.x)
            s[i] = i < n ? tmp[beg + i] : ~uint64_t(0);
        __syncthreads();
        for (uint32_t k = 2; k <= N; k <<= 1)
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
                    const uint32_t ixj = i ^ j;
                    if (ixj > i) {
                        const uint64_t a = s[i], c = s[ixj];
                        const bool up = (i & k) == 0;
                        if ((a > c) == up) {
                            s[i] = c;
                            s[ixj] = a;
                        }
                    }
                }
                __syncthreads();
            }
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) perm[beg + i] = uint32_t(s[i]);
    } else {
        // Oversized brick: rank by counting (keys are unique).
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const uint64_t v = tmp[beg + i];
            uint32_t r = 0;
            for (uint32_t q = 0; q < n; ++q) r += tmp[beg + q] < v;
            perm[beg + r] = uint32_t(v);
        }
    }
}

}  // namespace pic

using namespace pic;

extern "C" int64_t pic_bin_count(const int32_t ncell[3], const int32_t brick[3])
{
    if (!ncell || !brick) return PIC_ERR_ARG;
    int64_t nb = 1, bvol = 1;
    for (int d = 0; d < 3; ++d) {
        if (ncell[d] <= 0 || brick[d] <= 0) return PIC_ERR_ARG;
        nb *= (ncell[d] + brick[d] - 1) / brick[d];
        bvol *= brick[d];
    }
    if (nb * bvol >= int64_t(UINT32_MAX) || nb >= int64_t(INT32_MAX)) return PIC_ERR_CAPACITY;
    return nb;
}

extern "C" int pic_bin_sort(pic_context* ctx, int dtype, const pic_mesh* mesh, int64_t np,
                            const pic_pfield* x, const int32_t cell_lo[3], const int32_t ncell[3],
                            const int32_t brick[3], uint32_t* keys, uint32_t* perm,
                            uint32_t* bin_start)
{
    if (!ctx || !mesh || !(mesh->h > 0.0) || (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 ||
        np >= int64_t(UINT32_MAX) || !cell_lo || !ncell || !brick || !bin_start)
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !keys || !perm)) return PIC_ERR_ARG;
    const int64_t nb = pic_bin_count(ncell, brick);
    if (nb < 0) return (int)nb;
    KeyGeom g;
    g.bvol = 1;
    for (int d = 0; d < 3; ++d) {
        g.low[d] = mesh->low[d];
        g.cell_lo[d] = cell_lo[d];
        g.ncell[d] = ncell[d];
        g.brick[d] = brick[d];
        g.nb[d] = (ncell[d] + brick[d] - 1) / brick[d];
        g.bvol *= brick[d];
    }
    g.rdx = 1.0 / mesh->h;
    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;

    // scratch: count[nb+1] | cursor[nb] | scan tmp | tmp pairs (8-aligned)
    const int64_t scan_words = scan_tmp_words(nb + 1);
    int64_t words = (nb + 1) + nb + scan_words;
    words = (words + 1) & ~int64_t(1);
    const size_t bytes = size_t(words) * 4 + size_t(np) * 8;
    uint32_t* sc = static_cast<uint32_t*>(scratch(ctx, bytes));
    if (!sc) return PIC_ERR_ALLOC;
    uint32_t* count = sc;
    uint32_t* cursor = count + nb + 1;
    uint32_t* scan_tmp = cursor + nb;
    uint64_t* tmp = reinterpret_cast<uint64_t*>(sc + words);
    hipStream_t st = ctx->stream;

    TimedLaunch tl(ctx, "bin_sort");
    const unsigned fb = (unsigned)std::min<int64_t>((2 * nb + 1 + 255) / 256, 4096);
    k_fill_u32<<<fb, 256, 0, st>>>(count, 2 * nb + 1, 0u);
    if (np > 0) {
        const unsigned b = blocks_for(np, 256);
        if (dtype == PIC_F64)
            k_cell_keys<double><<<b, 256, 0, st>>>(np, make_pfield<double>(*x), g, keys, count,
                                                    ctx->d_err);
        else
            k_cell_keys<float><<<b, 256, 0, st>>>(np, make_pfield<float>(*x), g, keys, count,
                                                   ctx->d_err);
    }
    exclusive_scan(st, count, bin_start, nb + 1, scan_tmp);
    if (np > 0) {
        k_bin_scatter<<<blocks_for(np, 256), 256, 0, st>>>(np, keys, g.bvol, bin_start, cursor,
                                                            tmp);
        k_brick_sort<<<(unsigned)nb, 256, SORT_LDS_CAP * sizeof(uint64_t), st>>>(bin_start, tmp,
                                                                               perm);
    }
    return launch_status();
}
