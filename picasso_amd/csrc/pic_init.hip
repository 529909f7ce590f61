// Particle initialisation on the GPU: Picasso's initializeParticles
// (Picasso_ParticleInit.hpp:121-248, 279-400) fills every owned cell with
// particles, calls the application's creation functor on each and compacts
// away the ones it rejects (filterEmpties, :38-91).  Here the fill is one
// kernel writing positions and volumes straight into the caller's particle
// fields; the functor's part stays with the application (any GPU code that
// turns the positions into a keep mask and fills its own fields), and
// pic_select + pic_permute do the compaction, stably.
//
// Particle order is the reference's: cell_id * n + local, with
// cell_id = i + ncell_i (j + ncell_j k) over the owned cells (i fastest)
// and, for the uniform lattice, local = ip + d (jp + d kp).
#include "pic_common.hpp"
#include "pic_internal.hpp"

namespace pic {

// Counter-based random numbers: SplitMix64 of (seed, particle, dimension),
// top 53 bits as a double in [0, 1).  Deterministic for a seed and
// independent of the launch shape.  (The reference draws from a Kokkos
// XorShift64 pool seeded per rank; only the distribution is specified.)
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t pid, int d)
{
    uint64_t z = seed + (pid * 3u + uint64_t(d) + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_init_particles(int mode, int64_t np, int npc, int ppd,
                                                        double low0, double low1, double low2,
                                                        double h, int32_t lo0, int32_t lo1,
                                                        int32_t lo2, int32_t n0, int32_t n1,
                                                        uint64_t seed, PField<T> x, PField<T> vol)
{
#pragma clang fp contract(off)
    const double low[3] = {low0, low1, low2};
    const int32_t lo[3] = {lo0, lo1, lo2};
    for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < np;
         p += int64_t(gridDim.x) * blockDim.x) {
        const int64_t cell = p / npc;
        const int local = int(p - cell * npc);
        const int32_t own[3] = {int32_t(cell % n0), int32_t((cell / n0) % n1),
                                int32_t(cell / (int64_t(n0) * n1))};
        // Low and high nodes of the cell, as LocalMesh::coordinates gives
        // them: low corner + index * h.
        double lc[3], hc[3];
        for (int d = 0; d < 3; ++d) {
            lc[d] = low[d] + double(lo[d] + own[d]) * h;
            hc[d] = low[d] + double(lo[d] + own[d] + 1) * h;
        }
        double px[3];
        if (mode == PIC_INIT_UNIFORM) {
            const int ip[3] = {local % ppd, (local / ppd) % ppd, local / (ppd * ppd)};
            for (int d = 0; d < 3; ++d) {
                const double sp = (hc[d] - lc[d]) / double(ppd);
                px[d] = 0.5 * sp + double(ip[d]) * sp + lc[d];
            }
        } else {
            for (int d = 0; d < 3; ++d)
                px[d] = lc[d] + (hc[d] - lc[d]) * uniform01(seed, uint64_t(p), d);
        }
        const int64_t xb = x.base(p);
        for (int d = 0; d < 3; ++d) x.at(xb, d) = T(px[d]);
        if (vol.data) vol.at(vol.base(p), 0) = T(h * h * h / double(npc));
    }
}

}  // namespace pic

using namespace pic;

extern "C" int64_t pic_init_particles(pic_context* ctx, int mode, int dtype,
                                      const pic_mesh* mesh, const int32_t cell_lo[3],
                                      const int32_t ncell[3], int ppc, uint64_t seed,
                                      int64_t capacity, const pic_pfield* x,
                                      const pic_pfield* volume)
{
    Region region("Picasso::initializeParticles");
    if (!ctx || !mesh || !(mesh->h > 0.0) || !cell_lo || !ncell ||
        (mode != PIC_INIT_UNIFORM && mode != PIC_INIT_RANDOM) ||
        (dtype != PIC_F64 && dtype != PIC_F32) || ppc < 1 || capacity < 0)
        return PIC_ERR_ARG;
    int64_t cells = 1;
    for (int d = 0; d < 3; ++d) {
        if (ncell[d] < 1) return PIC_ERR_ARG;
        cells *= ncell[d];
    }
    const int64_t npc = mode == PIC_INIT_UNIFORM ? int64_t(ppc) * ppc * ppc : int64_t(ppc);
    if (npc > INT32_MAX || cells > INT64_MAX / npc) return PIC_ERR_CAPACITY;
    const int64_t np = cells * npc;
    if (np > capacity) return PIC_ERR_CAPACITY;
    if (!valid_pfield(x) || (volume && !valid_pfield(volume))) return PIC_ERR_ARG;
    DeviceGuard dg(ctx->device);
    if (!dg.ok) return PIC_ERR_HIP;
    TimedLaunch tl(ctx, "init_particles");
    const unsigned b = (unsigned)std::min<int64_t>(blocks_for(np, 256), 65536);
    auto go = [&](auto tag) {
        using T = decltype(tag);
        const PField<T> vf = volume ? make_pfield<T>(*volume) : PField<T>{};
        k_init_particles<T><<<b, 256, 0, ctx->stream>>>(
            mode, np, int(npc), ppc, mesh->low[0], mesh->low[1], mesh->low[2], mesh->h,
            cell_lo[0], cell_lo[1], cell_lo[2], ncell[0], ncell[1], seed, make_pfield<T>(*x), vf);
    };
    if (dtype == PIC_F64) go(double{});
    else go(float{});
    const int rc = launch_status();
    return rc < 0 ? rc : np;
}
