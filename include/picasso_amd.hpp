// picasso_amd.hpp -- header-only C++ facade over the C ABI in picasso_amd.h
// that keeps Picasso's argument order for the transfer path:
//
//   Picasso::APIC::p2g(m_p, c_p, m_i, mu_i, sd)          (Picasso_APIC.hpp:55-65)
//   Picasso::APIC::g2p(u_i, c_p, sd)                     (Picasso_APIC.hpp:407-413)
//   Picasso::PolyPIC::p2g(m_p, u_p, c_p, mu_i, m_i, dt, sd) (Picasso_PolyPIC.hpp:39-50)
//   Picasso::PolyPIC::g2p(u_i, c_p, sd)                  (Picasso_PolyPIC.hpp:235-243)
//   Picasso::G2P::value / gradient / divergence, Picasso::P2G::value /
//   gradient / divergence   (Picasso_ParticleInterpolation.hpp:110-279)
//   Picasso::initializeParticles(InitUniform / InitRandom, ...)
//                           (Picasso_ParticleInit.hpp:121-248, 279-400)
//   GridOperator's gather / scatter-sum halos on one GPU, periodic
//                           (Picasso_GridOperator.hpp:203-215, 465-491)
//
// The reference calls these once per particle inside a Kokkos kernel; here
// each call covers every particle of the spline data at once (the batch form
// the C ABI exports).  `sd` is the analogue of createSpline(Location,
// InterpolationOrder, local_mesh, x) (Picasso_ParticleInterpolation.hpp:
// 88-105): it records where the positions live, the spline order and the
// entity, and the kernels evaluate the weights themselves.  Errors are
// thrown as picasso_amd::Error carrying the pic_status; device-side errors
// surface at Context::synchronize().
#ifndef PICASSO_AMD_HPP
#define PICASSO_AMD_HPP

#include <exception>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "picasso_amd.h"

namespace picasso_amd {

struct Error : std::runtime_error {
    int status;
    explicit Error(int s)
        : std::runtime_error(std::string("picasso_amd: ") + pic_status_string(s)), status(s)
    {
    }
};

inline void check(int status)
{
    if (status != PIC_OK) throw Error(status);
}

// Owns a pic_context (device, stream binding, error word, scratch).
class Context {
  public:
    explicit Context(int device = 0) { check(pic_context_create(&ctx_, device)); }
    ~Context()
    {
        if (ctx_) pic_context_destroy(ctx_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    void set_stream(void* hip_stream) { check(pic_set_stream(ctx_, hip_stream)); }
    void synchronize() { check(pic_synchronize(ctx_)); }
    pic_context* get() const { return ctx_; }

  private:
    pic_context* ctx_ = nullptr;
};

// Particle field descriptors (see pic_pfield).
inline pic_pfield soa(void* data, int64_t np) { return pic_pfield{data, 0, 1, np}; }
inline pic_pfield aos(void* data, int ncomp) { return pic_pfield{data, 0, ncomp, 1}; }
// Cabana AoSoA slice: vector length 2^vlen_log2, tile of tile_stride
// scalars, data pointing at the member's first scalar
// (Picasso_ParticleList.hpp:90-138).
inline pic_pfield aosoa(void* data, int vlen_log2, int64_t tile_stride)
{
    return pic_pfield{data, vlen_log2, tile_stride, int64_t(1) << vlen_log2};
}

// Dense grid window [lo, lo+n) with the component index fastest.
inline pic_grid dense_grid(void* data, const int lo[3], const int n[3], int ncomp)
{
    pic_grid g;
    g.data = data;
    for (int d = 0; d < 3; ++d) {
        g.lo[d] = lo[d];
        g.n[d] = n[d];
    }
    g.cstride = 1;
    g.stride[2] = ncomp;
    g.stride[1] = int64_t(n[2]) * ncomp;
    g.stride[0] = int64_t(n[1]) * n[2] * ncomp;
    return g;
}

// Batch spline data: createSpline(Location, InterpolationOrder<order>,
// local_mesh, x) for particles [0, np).  bins (from pic_bin_sort) selects
// the brick kernels when the binning matches the spline's knot lattice.
struct SplineData {
    int order;
    int entity;
    int dtype;
    pic_mesh mesh;
    int64_t np;
    pic_pfield x;
    const pic_bins* bins;
};

inline SplineData createSpline(int entity, int order, int dtype, const pic_mesh& mesh,
                               int64_t np, const pic_pfield& x,
                               const pic_bins* bins = nullptr)
{
    return SplineData{order, entity, dtype, mesh, np, x, bins};
}

namespace APIC {

inline void p2g(pic_context* ctx, const pic_pfield& m_p, const pic_pfield& c_p,
                const pic_grid& m_i, const pic_grid& mu_i, const SplineData& sd,
                uint32_t flags = 0)
{
    check(pic_apic_p2g(ctx, sd.order, sd.entity, sd.dtype, &sd.mesh, sd.np, &sd.x, &m_p, &c_p,
                       &m_i, &mu_i, sd.bins, flags));
}

inline void g2p(pic_context* ctx, const pic_grid& u_i, const pic_pfield& c_p,
                const SplineData& sd)
{
    check(pic_apic_g2p(ctx, sd.order, sd.entity, sd.dtype, &sd.mesh, sd.np, &sd.x, &u_i, &c_p,
                       sd.bins));
}

// g2p of u_i = mu_i / m_i formed on the fly (grid normalisation fused in).
inline void g2p_normalized(pic_context* ctx, const pic_grid& m_i, const pic_grid& mu_i,
                           const pic_pfield& c_p, const SplineData& sd)
{
    check(pic_apic_g2p_normalized(ctx, sd.order, sd.entity, sd.dtype, &sd.mesh, sd.np, &sd.x,
                                  &m_i, &mu_i, &c_p, sd.bins));
}

// One step in one call (pic_apic_step): the bin sort of sd's particles by
// geom into keys / perm / bin_start / active, p2g into m_i / mu_i with the
// grid zeroed first and g2p of mu_i / m_i into c_p, with fewer launches than
// binSort + p2g + g2p_normalized.  sd.bins is not used; the bins the step
// wrote are returned.
inline pic_bins step(pic_context* ctx, const pic_pfield& m_p, const pic_pfield& c_p,
                     const pic_grid& m_i, const pic_grid& mu_i, const SplineData& sd,
                     const pic_binning& geom, uint32_t* keys, uint32_t* perm,
                     uint32_t* bin_start, uint32_t* active)
{
    check(pic_apic_step(ctx, sd.order, sd.entity, sd.dtype, &sd.mesh, sd.np, &sd.x, &m_p, &c_p,
                        &m_i, &mu_i, &geom, keys, perm, bin_start, active));
    return pic_bins{perm, keys, bin_start, active, geom};
}

// g2p fused with the velocity gradient and the FP64 update of F (and its
// determinant and inverse).  m_i non-null: u_i is the grid momentum.
inline void g2p_update_F(pic_context* ctx, const pic_grid* m_i, const pic_grid& u_i,
                         const pic_pfield& c_p, double dt, const pic_pfield& F,
                         const pic_pfield& det, const pic_pfield& Finv, const SplineData& sd)
{
    check(pic_apic_g2p_update_F(ctx, sd.order, sd.entity, sd.dtype, &sd.mesh, sd.np, &sd.x, m_i,
                                &u_i, &c_p, dt, &F, &det, &Finv, sd.bins));
}

}  // namespace APIC

// Generic transfers, Picasso::G2P / P2G (Picasso_ParticleInterpolation.hpp:
// 110-279): ncomp counts the input's components (see pic_g2p_interp).
namespace G2P {

inline void value(pic_context* ctx, const SplineData& sd, const pic_grid& field, int ncomp,
                  const pic_pfield& result)
{
    check(pic_g2p_interp(ctx, PIC_INTERP_VALUE, sd.order, sd.entity, sd.dtype, &sd.mesh, sd.np,
                         &sd.x, &field, ncomp, &result, sd.bins));
}

inline void gradient(pic_context* ctx, const SplineData& sd, const pic_grid& field, int ncomp,
                     const pic_pfield& result)
{
    check(pic_g2p_interp(ctx, PIC_INTERP_GRADIENT, sd.order, sd.entity, sd.dtype, &sd.mesh,
                         sd.np, &sd.x, &field, ncomp, &result, sd.bins));
}

inline void divergence(pic_context* ctx, const SplineData& sd, const pic_grid& field,
                       const pic_pfield& result)
{
    check(pic_g2p_interp(ctx, PIC_INTERP_DIVERGENCE, sd.order, sd.entity, sd.dtype, &sd.mesh,
                         sd.np, &sd.x, &field, 3, &result, sd.bins));
}

}  // namespace G2P

namespace P2G {

inline void value(pic_context* ctx, const SplineData& sd, const pic_pfield& value, int ncomp,
                  const pic_grid& field, uint32_t flags = 0)
{
    check(pic_p2g_interp(ctx, PIC_INTERP_VALUE, sd.order, sd.entity, sd.dtype, &sd.mesh, sd.np,
                         &sd.x, &value, ncomp, &field, flags, sd.bins));
}

inline void gradient(pic_context* ctx, const SplineData& sd, const pic_pfield& value,
                     const pic_grid& field, uint32_t flags = 0)
{
    check(pic_p2g_interp(ctx, PIC_INTERP_GRADIENT, sd.order, sd.entity, sd.dtype, &sd.mesh,
                         sd.np, &sd.x, &value, 1, &field, flags, sd.bins));
}

// Vector values (ncomp 3) or row-major tensor values (ncomp 9).
inline void divergence(pic_context* ctx, const SplineData& sd, const pic_pfield& value,
                       int ncomp, const pic_grid& field, uint32_t flags = 0)
{
    check(pic_p2g_interp(ctx, PIC_INTERP_DIVERGENCE, sd.order, sd.entity, sd.dtype, &sd.mesh,
                         sd.np, &sd.x, &value, ncomp, &field, flags, sd.bins));
}

}  // namespace P2G

namespace PolyPIC {

inline void p2g(pic_context* ctx, const pic_pfield& m_p, const pic_pfield& u_p,
                const pic_pfield& c_p, const pic_grid& mu_i, const pic_grid& m_i, double dt,
                const SplineData& sd, int ncomp = 3, uint32_t flags = 0)
{
    check(pic_polypic_p2g(ctx, sd.order, sd.entity, sd.dtype, &sd.mesh, sd.np, ncomp, &sd.x,
                          &m_p, &u_p, &c_p, &mu_i, &m_i, dt, sd.bins, flags));
}

inline void g2p(pic_context* ctx, const pic_grid& u_i, const pic_pfield& c_p,
                const SplineData& sd, int ncomp = 3)
{
    check(pic_polypic_g2p(ctx, sd.order, sd.entity, sd.dtype, &sd.mesh, sd.np, ncomp, &sd.x,
                          &u_i, &c_p, sd.bins));
}

}  // namespace PolyPIC

// Particle initialisation without the creation functor: positions (and
// volumes) of every particle of the cells [cell_lo, cell_lo + ncell); the
// functor's choice then goes through select + pic_permute.  Returns the
// number of particles written.
struct InitUniform {};
struct InitRandom {};

inline int64_t initializeParticles(pic_context* ctx, InitUniform, int particles_per_cell_dim,
                                   int dtype, const pic_mesh& mesh, const int32_t cell_lo[3],
                                   const int32_t ncell[3], int64_t capacity,
                                   const pic_pfield& x, const pic_pfield* volume = nullptr)
{
    const int64_t n = pic_init_particles(ctx, PIC_INIT_UNIFORM, dtype, &mesh, cell_lo, ncell,
                                         particles_per_cell_dim, 0, capacity, &x, volume);
    if (n < 0) check(int(n));
    return n;
}

inline int64_t initializeParticles(pic_context* ctx, InitRandom, int particles_per_cell,
                                   int dtype, const pic_mesh& mesh, const int32_t cell_lo[3],
                                   const int32_t ncell[3], int64_t capacity,
                                   const pic_pfield& x, const pic_pfield* volume = nullptr,
                                   uint64_t seed = 0)
{
    const int64_t n = pic_init_particles(ctx, PIC_INIT_RANDOM, dtype, &mesh, cell_lo, ncell,
                                         particles_per_cell, seed, capacity, &x, volume);
    if (n < 0) check(int(n));
    return n;
}

// Stable counting sort by bin key (pic_bin_sort); returns the pic_bins that
// createSpline takes to select the brick kernels.  keys and perm hold np
// words, bin_start and active pic_bin_count(&geom) + 1.
inline pic_bins binSort(pic_context* ctx, int dtype, const pic_mesh& mesh, int64_t np,
                        const pic_pfield& x, const pic_binning& geom, uint32_t* keys,
                        uint32_t* perm, uint32_t* bin_start, uint32_t* active)
{
    check(pic_bin_sort(ctx, dtype, &mesh, np, &x, &geom, keys, perm, bin_start, active));
    return pic_bins{perm, keys, bin_start, active, geom};
}

// dst(s, c) = src(perm[s], c) for each of nfields fields (pic_permute).
inline void permute(pic_context* ctx, int dtype, int64_t np, const uint32_t* perm, int nfields,
                    const pic_pfield* src, const pic_pfield* dst, const int* ncomp)
{
    check(pic_permute(ctx, dtype, np, perm, nfields, src, dst, ncomp));
}

// kept[0] = number of kept particles, kept[1..] their indices in order.
inline void select(pic_context* ctx, int64_t np, const uint8_t* keep, uint32_t* kept)
{
    check(pic_select(ctx, np, keep, kept));
}

// GridOperator::apply for particle operations on one GPU
// (Picasso_GridOperator.hpp:459-496): gather halos of the gather fields,
// zeroed scatter fields, op(ctx) issuing the transfers, scatter-sum halos.
class GridOperator {
  public:
    GridOperator(std::vector<pic_field_dep> gather, std::vector<pic_field_dep> scatter)
        : gather_(std::move(gather)), scatter_(std::move(scatter))
    {
    }
    template <class Op>
    void apply(pic_context* ctx, int dtype, Op&& op) const
    {
        struct Call {
            Op* op;
            std::exception_ptr err;
            static int run(pic_context* c, void* self)
            {
                Call* me = static_cast<Call*>(self);
                try {
                    (*me->op)(c);
                    return PIC_OK;
                } catch (...) {
                    me->err = std::current_exception();
                    return PIC_ERR_ARG;
                }
            }
        } call{&op, nullptr};
        const int rc = pic_grid_operator_apply(ctx, dtype, int(gather_.size()), gather_.data(),
                                               int(scatter_.size()), scatter_.data(), &Call::run,
                                               &call);
        if (call.err) std::rethrow_exception(call.err);
        check(rc);
    }

  private:
    std::vector<pic_field_dep> gather_, scatter_;
};

namespace Halo {

inline void scatterSum(pic_context* ctx, int dtype, const pic_grid& g, int ncomp,
                       const int32_t own_lo[3], const int32_t own_n[3],
                       const int32_t periodic[3])
{
    check(pic_grid_halo(ctx, PIC_HALO_SCATTER_SUM, dtype, &g, ncomp, own_lo, own_n, periodic));
}

inline void gather(pic_context* ctx, int dtype, const pic_grid& g, int ncomp,
                   const int32_t own_lo[3], const int32_t own_n[3], const int32_t periodic[3])
{
    check(pic_grid_halo(ctx, PIC_HALO_GATHER, dtype, &g, ncomp, own_lo, own_n, periodic));
}

}  // namespace Halo

}  // namespace picasso_amd

#endif
