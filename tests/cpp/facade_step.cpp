// One APIC step driven through the C++ facade (include/picasso_amd.hpp) on
// the GPU, as a Picasso application would call it, checked against the CPU
// oracle (test infrastructure, oracle/orpic.h).
//
// Particles live in one Cabana-style AoSoA of vector length 16 with members
// x (3), m (1) and c_p (12) (Picasso_ParticleList.hpp:90-138): 16^3 cells at
// 8 random particles per cell, FP64.  The step is bin sort -> APIC::p2g
// (zero first) -> APIC::g2p with the grid normalisation fused in, through
// the brick kernels; the oracle runs p2g -> normalise -> g2p on the same
// inputs.  The same step then runs again from the same inputs as one
// APIC::step call (pic_apic_step), into grids filled with garbage.  Prints
// the relative errors and exits 0 when every field of both runs is within
// 1e-12 (FP64; atomic summation order differs), 1 otherwise.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "orpic.h"
#include "picasso_amd.hpp"

namespace pa = picasso_amd;

namespace {

constexpr int N = 16, PPC = 8, HALO = 2, V = 16, Q = 16;  // Q scalars per particle

// SplitMix64 uniform in [0, 1).
struct Rng {
    uint64_t s;
    double next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return double(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

double rel_inf(const std::vector<double>& got, const std::vector<double>& ref)
{
    double num = 0.0, den = 0.0;
    for (size_t i = 0; i < ref.size(); ++i) {
        num = std::fmax(num, std::fabs(got[i] - ref[i]));
        den = std::fmax(den, std::fabs(ref[i]));
    }
    return den > 0.0 ? num / den : num;
}

bool ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    return e == hipSuccess;
}

}  // namespace

int main()
{
    const double h = 1.0 / N;
    const int64_t P = int64_t(N) * N * N * PPC;
    const int64_t ntile = (P + V - 1) / V;
    // Host AoS copies (oracle layout) and the packed AoSoA.
    std::vector<double> x(P * 3), m(P), c(P * 12), aosoa(ntile * V * Q, 0.0);
    Rng rng{0x5EED0C00ull};
    int64_t p = 0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k)
                for (int q = 0; q < PPC; ++q, ++p) {
                    const int cell[3] = {i, j, k};
                    for (int d = 0; d < 3; ++d) x[p * 3 + d] = (cell[d] + rng.next()) * h;
                    m[p] = h * h * h / PPC;
                    for (int e = 0; e < 12; ++e)
                        c[p * 12 + e] = (e < 3 ? 1.0 : 0.1) * (2.0 * rng.next() - 1.0);
                }
    // Member offsets inside a particle's struct: x 0..2, m 3, c_p 4..15.
    auto at = [&](int64_t pp, int s) -> double& {
        return aosoa[(pp / V) * V * Q + int64_t(s) * V + pp % V];
    };
    for (p = 0; p < P; ++p) {
        for (int d = 0; d < 3; ++d) at(p, d) = x[p * 3 + d];
        at(p, 3) = m[p];
        for (int e = 0; e < 12; ++e) at(p, 4 + e) = c[p * 12 + e];
    }

    const int nn = N + 1 + 2 * HALO;
    const int lo[3] = {0, 0, 0}, n[3] = {nn, nn, nn};
    const int64_t nodes = int64_t(nn) * nn * nn;
    const pic_mesh mesh{{-HALO * h, -HALO * h, -HALO * h}, h};
    // Bins: node-centred dual cells (the quadratic knot base), 4^3 bricks.
    const pic_binning geom{{HALO, HALO, HALO}, {N + 1, N + 1, N + 1}, {4, 4, 4}, 1};
    const int64_t nb = pic_bin_count(&geom);

    double *d_p = nullptr, *d_m = nullptr, *d_mu = nullptr;
    uint32_t* d_u32 = nullptr;
    if (!ok(hipMalloc(&d_p, aosoa.size() * sizeof(double)), "hipMalloc") ||
        !ok(hipMalloc(&d_m, nodes * sizeof(double)), "hipMalloc") ||
        !ok(hipMalloc(&d_mu, nodes * 3 * sizeof(double)), "hipMalloc") ||
        !ok(hipMalloc(&d_u32, (2 * P + 2 * (nb + 1)) * sizeof(uint32_t)), "hipMalloc") ||
        !ok(hipMemcpy(d_p, aosoa.data(), aosoa.size() * sizeof(double), hipMemcpyHostToDevice),
            "hipMemcpy"))
        return 1;
    const std::vector<double> aosoa0 = aosoa;
    std::vector<double> gm(nodes), gmu(nodes * 3), cg(P * 12);
    try {
        pa::Context ctx(0);
        const pic_pfield px = pa::aosoa(d_p, 4, V * Q);
        const pic_pfield pm = pa::aosoa(d_p + 3 * V, 4, V * Q);
        const pic_pfield pc = pa::aosoa(d_p + 4 * V, 4, V * Q);
        const pic_grid g_m = pa::dense_grid(d_m, lo, n, 1);
        const pic_grid g_mu = pa::dense_grid(d_mu, lo, n, 3);
        const pic_bins bins = pa::binSort(ctx.get(), PIC_F64, mesh, P, px, geom, d_u32,
                                          d_u32 + P, d_u32 + 2 * P, d_u32 + 2 * P + nb + 1);
        const pa::SplineData sd = pa::createSpline(PIC_NODE, 2, PIC_F64, mesh, P, px, &bins);
        pa::APIC::p2g(ctx.get(), pm, pc, g_m, g_mu, sd, PIC_P2G_ZERO_FIRST);
        pa::APIC::g2p_normalized(ctx.get(), g_m, g_mu, pc, sd);
        ctx.synchronize();
    } catch (const pa::Error& e) {
        std::fprintf(stderr, "facade call failed: %s\n", e.what());
        return 1;
    }
    if (!ok(hipMemcpy(gm.data(), d_m, nodes * sizeof(double), hipMemcpyDeviceToHost), "copy") ||
        !ok(hipMemcpy(gmu.data(), d_mu, nodes * 3 * sizeof(double), hipMemcpyDeviceToHost),
            "copy") ||
        !ok(hipMemcpy(aosoa.data(), d_p, aosoa.size() * sizeof(double), hipMemcpyDeviceToHost),
            "copy"))
        return 1;
    for (p = 0; p < P; ++p)
        for (int e = 0; e < 12; ++e) cg[p * 12 + e] = at(p, 4 + e);

    std::vector<double> rm(nodes, 0.0), rmu(nodes * 3, 0.0), ru(nodes * 3), rc(P * 12);
    const int32_t lo32[3] = {0, 0, 0}, n32[3] = {nn, nn, nn};
    if (orc_apic_p2g(2, ORC_NODE, mesh.low, h, lo32, n32, P, x.data(), m.data(), c.data(),
                     rm.data(), rmu.data()) ||
        orc_grid_normalize(nodes, 3, rm.data(), rmu.data(), ru.data()) ||
        orc_apic_g2p(2, ORC_NODE, mesh.low, h, lo32, n32, P, x.data(), ru.data(), rc.data())) {
        std::fprintf(stderr, "oracle failed\n");
        return 1;
    }
    const double em = rel_inf(gm, rm), emu = rel_inf(gmu, rmu), ec = rel_inf(cg, rc);
    std::printf("facade APIC step, %lld particles in a V=%d AoSoA: rel err mass %.3e "
                "momentum %.3e c_p %.3e\n",
                (long long)P, V, em, emu, ec);

    // The same step as one call, from the same inputs, into garbage grids.
    aosoa = aosoa0;
    if (!ok(hipMemcpy(d_p, aosoa.data(), aosoa.size() * sizeof(double), hipMemcpyHostToDevice),
            "hipMemcpy") ||
        !ok(hipMemset(d_m, 0x41, nodes * sizeof(double)), "hipMemset") ||
        !ok(hipMemset(d_mu, 0x41, nodes * 3 * sizeof(double)), "hipMemset"))
        return 1;
    try {
        pa::Context ctx(0);
        const pic_pfield px = pa::aosoa(d_p, 4, V * Q);
        const pic_pfield pm = pa::aosoa(d_p + 3 * V, 4, V * Q);
        const pic_pfield pc = pa::aosoa(d_p + 4 * V, 4, V * Q);
        const pic_grid g_m = pa::dense_grid(d_m, lo, n, 1);
        const pic_grid g_mu = pa::dense_grid(d_mu, lo, n, 3);
        const pa::SplineData sd = pa::createSpline(PIC_NODE, 2, PIC_F64, mesh, P, px);
        (void)pa::APIC::step(ctx.get(), pm, pc, g_m, g_mu, sd, geom, d_u32, d_u32 + P,
                             d_u32 + 2 * P, d_u32 + 2 * P + nb + 1);
        ctx.synchronize();
    } catch (const pa::Error& e) {
        std::fprintf(stderr, "facade step call failed: %s\n", e.what());
        return 1;
    }
    if (!ok(hipMemcpy(gm.data(), d_m, nodes * sizeof(double), hipMemcpyDeviceToHost), "copy") ||
        !ok(hipMemcpy(gmu.data(), d_mu, nodes * 3 * sizeof(double), hipMemcpyDeviceToHost),
            "copy") ||
        !ok(hipMemcpy(aosoa.data(), d_p, aosoa.size() * sizeof(double), hipMemcpyDeviceToHost),
            "copy"))
        return 1;
    for (p = 0; p < P; ++p)
        for (int e = 0; e < 12; ++e) cg[p * 12 + e] = at(p, 4 + e);
    const double sm = rel_inf(gm, rm), smu = rel_inf(gmu, rmu), sc = rel_inf(cg, rc);
    std::printf("facade APIC::step (one call): rel err mass %.3e momentum %.3e c_p %.3e\n", sm,
                smu, sc);
    (void)hipFree(d_p);
    (void)hipFree(d_m);
    (void)hipFree(d_mu);
    (void)hipFree(d_u32);
    return (em < 1e-12 && emu < 1e-12 && ec < 1e-12 && sm < 1e-12 && smu < 1e-12 && sc < 1e-12)
               ? 0
               : 1;
}
