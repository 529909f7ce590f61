// Grid normalisation, the particle velocity gradient and the batched 3x3
// deformation-gradient update.
#include "pic_common.hpp"
#include "pic_internal.hpp"

namespace pic {

// ---------------------------------------------------------------------------
// Grid normalisation u = mu / m.

// Largest mass in the window (masses are >= 0, so the IEEE bit patterns of
// floats order like unsigned integers).
__global__ __launch_bounds__(256) void k_max_mass_f32(GView<float> m, unsigned* out)
{
    const int64_t n12 = int64_t(m.n[1]) * m.n[2];
    const int64_t total = int64_t(m.n[0]) * n12;
    float best = 0.f;
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int64_t i = t / n12, r = t % n12, j = r / m.n[2], k = r % m.n[2];
        best = fmaxf(best, m.data[i * m.s[0] + j * m.s[1] + k * m.s[2]]);
    }
    // wave reduction, then one atomic per wave
    for (int o = 32; o > 0; o >>= 1) best = fmaxf(best, __shfl_xor(best, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(best));
}

__global__ void k_set_u32(unsigned* p, unsigned v) { *p = v; }

template <typename T>
__global__ __launch_bounds__(256) void k_normalize(GView<T> m, GView<T> mu, GView<T> u, int ncomp,
                                                   const unsigned* mmax, float rel_eps)
{
    const int64_t n12 = int64_t(m.n[1]) * m.n[2];
    const int64_t total = int64_t(m.n[0]) * n12;
    T thr = T(0);
    if constexpr (std::is_same<T, float>::value) {
        if (mmax) thr = rel_eps * __uint_as_float(*mmax);
    }
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int i = int(t / n12), r = int(t % n12), j = r / m.n[2], k = r % m.n[2];
        const T mm = m.data[int64_t(i) * m.s[0] + int64_t(j) * m.s[1] + int64_t(k) * m.s[2]];
        const int64_t om = int64_t(i) * mu.s[0] + int64_t(j) * mu.s[1] + int64_t(k) * mu.s[2];
        const int64_t ou = int64_t(i) * u.s[0] + int64_t(j) * u.s[1] + int64_t(k) * u.s[2];
        for (int d = 0; d < ncomp; ++d) {
            const T q = mu.data[om + d * mu.cs];
            T v;
            if (!(mm > T(0))) {
                v = T(0);
            } else if (std::is_same<T, float>::value && mm < thr) {
                v = (T)((double)q / (double)mm);
            } else {
                v = q / mm;
            }
            u.data[ou + d * u.cs] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Velocity gradient (G2P::gradient): L(a,b) = sum_n u_n(a) dw_n/dx_b.
template <typename T>
struct GraduArgs {
    Lattice lat;
    int64_t np;
    const uint32_t* perm;
    PField<T> x, L;
    GView<T> gu;
    int* err;
};

template <int ORDER, typename T>
__global__ __launch_bounds__(256) void k_g2p_gradu(GraduArgs<T> a)
{
    constexpr int NK = Knots<ORDER>::NK;
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= a.np) return;
    const int64_t p = a.perm ? int64_t(a.perm[s]) : s;
    const int64_t xb = a.x.base(p);
    T xp[3] = {a.x.at(xb, 0), a.x.at(xb, 1), a.x.at(xb, 2)};
    Spline<ORDER, T> sd;
    eval_spline<ORDER, T, true, false>(a.lat, xp, sd);
    if (!stencil_inside(sd, a.gu)) {
        atomicOr(a.err, ERRBIT_BOUNDS);
        return;
    }
    T L[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) L[q] = T(0);
#pragma unroll
    for (int i = 0; i < NK; ++i)
#pragma unroll
        for (int j = 0; j < NK; ++j)
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                const T gw[3] = {sd.g[0][i] * sd.w[1][j] * sd.w[2][k],
                                 sd.w[0][i] * sd.g[1][j] * sd.w[2][k],
                                 sd.w[0][i] * sd.w[1][j] * sd.g[2][k]};
                const int64_t off = a.gu.off(sd.s0[0] + i, sd.s0[1] + j, sd.s0[2] + k);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T ue = a.gu.data[off + c * a.gu.cs];
#pragma unroll
                    for (int b = 0; b < 3; ++b) L[c * 3 + b] += ue * gw[b];
                }
            }
    const int64_t lb = a.L.base(p);
#pragma unroll
    for (int q = 0; q < 9; ++q) a.L.at(lb, q) = L[q];
}

// ---------------------------------------------------------------------------
// F <- (I + dt L) F; det(F); F^{-1}, all in FP64.
template <typename TL>
__global__ __launch_bounds__(256) void k_update_F(int64_t np, PField<TL> L, double dt,
                                                  PField<double> F, PField<double> det,
                                                  PField<double> Finv)
{
    const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int64_t lb = L.base(p), fb = F.base(p);
    double Lp[9], Fp[9], inv[9], dd;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        Lp[q] = (double)L.at(lb, q);
        Fp[q] = F.at(fb, q);
    }
    update_F(Lp, dt, Fp, dd, inv);
    const int64_t ib = Finv.base(p);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        F.at(fb, q) = Fp[q];
        Finv.at(ib, q) = inv[q];
    }
    det.at(det.base(p), 0) = dd;
}

}  // namespace pic

using namespace pic;

extern "C" int pic_grid_normalize(pic_context* ctx, int dtype, int ncomp, const pic_grid* m_i,
                                  const pic_grid* mu_i, const pic_grid* u_i, double rel_mass_eps)
{
    Region region("picasso_amd::grid_normalize");
    if (!ctx || (dtype != PIC_F64 && dtype != PIC_F32) || ncomp < 1 || ncomp > 3 ||
        !valid_grid(m_i) || !valid_grid(mu_i) || !valid_grid(u_i) || rel_mass_eps < 0.0)
        return PIC_ERR_ARG;
    for (int d = 0; d < 3; ++d)
        if (mu_i->n[d] != m_i->n[d] || u_i->n[d] != m_i->n[d] || mu_i->lo[d] != m_i->lo[d] ||
            u_i->lo[d] != m_i->lo[d])
            return PIC_ERR_ARG;
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    const int64_t total = int64_t(m_i->n[0]) * m_i->n[1] * m_i->n[2];
    const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 8192);
    if (dtype == PIC_F64) {
        TimedLaunch tl(ctx, "normalize");
        k_normalize<double><<<blocks, 256, 0, ctx->stream>>>(
            make_gview<double>(*m_i), make_gview<double>(*mu_i), make_gview<double>(*u_i), ncomp,
            nullptr, 0.f);
        return launch_status();
    }
    unsigned* mmax = nullptr;
    if (rel_mass_eps > 0.0) {
        mmax = static_cast<unsigned*>(scratch(ctx, sizeof(unsigned)));
        if (!mmax) return PIC_ERR_ALLOC;
        TimedLaunch tl(ctx, "max_mass");
        k_set_u32<<<1, 1, 0, ctx->stream>>>(mmax, 0u);
        k_max_mass_f32<<<std::min(blocks, 2048u), 256, 0, ctx->stream>>>(
            make_gview<float>(*m_i), mmax);
        if (launch_status()) return PIC_ERR_HIP;
    }
    TimedLaunch tl(ctx, "normalize");
    k_normalize<float><<<blocks, 256, 0, ctx->stream>>>(make_gview<float>(*m_i),
                                                        make_gview<float>(*mu_i),
                                                        make_gview<float>(*u_i), ncomp, mmax,
                                                        (float)rel_mass_eps);
    return launch_status();
}

extern "C" int pic_g2p_gradu(pic_context* ctx, int order, int entity, int dtype,
                             const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                             const pic_grid* u_i, const pic_pfield* gradu, const pic_bins* bins)
{
    Region region("picasso_amd::G2P::gradu");
    if (!ctx || !mesh || !(mesh->h > 0.0) || order < 1 || order > 3 ||
        (entity != PIC_NODE && entity != PIC_CELL) || (dtype != PIC_F64 && dtype != PIC_F32) ||
        np < 0 || np > int64_t(UINT32_MAX) || !valid_grid(u_i))
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !valid_pfield(gradu))) return PIC_ERR_ARG;
    if (bins && !bins->perm) return PIC_ERR_ARG;
    if (np == 0) return PIC_OK;
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    TimedLaunch tl(ctx, "g2p_gradu");
    const unsigned b = blocks_for(np, 256);
    auto go = [&](auto tag) {
        using T = decltype(tag);
        GraduArgs<T> a{make_lattice(*mesh, entity), np, bins ? bins->perm : nullptr,
                       make_pfield<T>(*x), make_pfield<T>(*gradu), make_gview<T>(*u_i),
                       ctx->d_err};
        if (order == 1) k_g2p_gradu<1, T><<<b, 256, 0, ctx->stream>>>(a);
        else if (order == 2) k_g2p_gradu<2, T><<<b, 256, 0, ctx->stream>>>(a);
        else k_g2p_gradu<3, T><<<b, 256, 0, ctx->stream>>>(a);
    };
    if (dtype == PIC_F64) go(double{});
    else go(float{});
    return launch_status();
}

extern "C" int pic_update_F(pic_context* ctx, int64_t np, int gradu_dtype, const pic_pfield* gradu,
                            double dt, const pic_pfield* F, const pic_pfield* det,
                            const pic_pfield* Finv)
{
    Region region("picasso_amd::update_F");
    if (!ctx || np < 0 || (gradu_dtype != PIC_F64 && gradu_dtype != PIC_F32)) return PIC_ERR_ARG;
    if (np == 0) return PIC_OK;
    if (!valid_pfield(gradu) || !valid_pfield(F) || !valid_pfield(det) || !valid_pfield(Finv))
        return PIC_ERR_ARG;
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    TimedLaunch tl(ctx, "update_F");
    const unsigned b = blocks_for(np, 256);
    if (gradu_dtype == PIC_F64)
        k_update_F<double><<<b, 256, 0, ctx->stream>>>(np, make_pfield<double>(*gradu), dt,
                                                       make_pfield<double>(*F),
                                                       make_pfield<double>(*det),
                                                       make_pfield<double>(*Finv));
    else
        k_update_F<float><<<b, 256, 0, ctx->stream>>>(np, make_pfield<float>(*gradu), dt,
                                                      make_pfield<double>(*F),
                                                      make_pfield<double>(*det),
                                                      make_pfield<double>(*Finv));
    return launch_status();
}
