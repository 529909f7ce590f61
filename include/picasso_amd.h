/* picasso_amd.h -- C ABI of the MI355X (gfx950) particle<->grid transfer
 * library.
 *
 * The reference (Picasso) exposes this path as header-only C++ device
 * functions called once per particle inside the caller's Kokkos kernel; there
 * is no FFI.  Per-particle calls cannot cross a shared-library boundary
 * efficiently, so each entry point here is the batch form of one reference
 * function: it runs that function for particles [0, np) on the context's HIP
 * stream.  Argument meaning and order follow the reference; the citations
 * name the function each entry point replaces.
 *
 * Conventions
 *  - Every call returns int: 0 (PIC_OK) or a negative pic_status.  Nothing
 *    throws across the ABI.  Calls are asynchronous on the context's stream;
 *    errors detected on the device (a stencil leaving its grid window, a
 *    particle outside the binned cells) are latched in the context and
 *    returned by pic_synchronize().
 *  - All buffers are caller-owned device memory.  Particle fields are
 *    described by pic_pfield (SoA, AoS and Cabana AoSoA all map onto it),
 *    grid fields by pic_grid (any i/j/k/component strides).
 *  - p2g accumulates (+=) like Kokkos::ScatterView::access() in the
 *    reference (Picasso_APIC.hpp:69-73); PIC_P2G_ZERO_FIRST zeroes the grid
 *    windows first, as GridOperator does before a scatter
 *    (Picasso_GridOperator.hpp:532-533).
 *  - g2p overwrites the particle field (Picasso_APIC.hpp:463-471,
 *    Picasso_PolyPIC.hpp:260), except a staggered PolyPIC g2p, which only
 *    rewrites the column it works on (Picasso_PolyPIC.hpp:328).
 *  - dtype PIC_F64: every field is double.  PIC_F32: every field is float;
 *    the position->knot mapping runs in FP64.  p2g through bins sums each
 *    bin's contributions in FP32 registers, adds the bin sums that land on a
 *    node in FP64, and applies one FP32 atomic per node and brick.
 *  - Brick kernels: p2g and g2p take the brick path when bins are given, the
 *    binning is aligned with the spline's knot base in every dimension
 *    (dual_d = 1 where the spline is quadratic on a node-aligned dimension
 *    or linear/cubic on a cell-centred one, else 0: dual = 1 for quadratic
 *    Node or linear/cubic Cell, 0 for linear/cubic Node or quadratic Cell,
 *    PIC_DUAL_DIMS | mask for the staggered Face/Edge lattices) and bricks
 *    are 4x4x4 (p2g) or 4x4x4 / 8x8x8 (g2p; PolyPIC 4x4x4).  APIC and
 *    PolyPIC take them on every entity lattice (PolyPIC collocated with
 *    three columns).  Otherwise the
 *    per-particle atomic scatter / gather kernels run.  Particles whose
 *    position no longer matches their bin go through the per-particle path,
 *    so results never depend on the bins being current.
 */
#ifndef PICASSO_AMD_H
#define PICASSO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PIC_API_VERSION 1

typedef enum pic_status {
    PIC_OK = 0,
    PIC_ERR_ARG = -1,         /* invalid argument (null, bad enum, bad size) */
    PIC_ERR_BOUNDS = -2,      /* a stencil or cell fell outside its window */
    PIC_ERR_HIP = -3,         /* a HIP runtime call failed */
    PIC_ERR_ALLOC = -4,       /* device allocation failed */
    PIC_ERR_UNSUPPORTED = -5, /* combination not implemented */
    PIC_ERR_CAPACITY = -6     /* a size exceeds a documented maximum */
} pic_status;

/* Grid entity kinds (Picasso_FieldTypes.hpp:125-168). */
typedef enum pic_entity {
    PIC_NODE = 0,
    PIC_CELL = 1,
    PIC_FACE_I = 2,
    PIC_FACE_J = 3,
    PIC_FACE_K = 4,
    PIC_EDGE_I = 5,
    PIC_EDGE_J = 6,
    PIC_EDGE_K = 7
} pic_entity;

typedef enum pic_dtype { PIC_F64 = 0, PIC_F32 = 1 } pic_dtype;

/* p2g flags.  PIC_P2G_ZERO_FIRST zeroes the grid windows first (as
 * GridOperator resets its scatter fields, Picasso_GridOperator.hpp:532-533).
 * PIC_P2G_DETERMINISTIC (pic_apic_p2g with bins the brick kernels take; no
 * reference counterpart: Kokkos' ScatterView sums in any order) makes the
 * result bit-identical from run to run for the same inputs and bins: the
 * bricks are flushed in eight passes, one per brick colour (the parities of
 * the brick coordinates), so no two bricks in flight share a node and every
 * node receives its per-brick sums in colour order without atomics; the
 * particles whose bins are stale are then applied one after another in slot
 * order.  Slower than the default (eight launches; stale particles
 * serially); without bins, or with bins the brick kernels cannot use, the
 * call returns PIC_ERR_UNSUPPORTED before doing anything.  More than 8192
 * stale particles are still applied, but in an unspecified order, and
 * pic_synchronize reports PIC_ERR_CAPACITY. */
enum { PIC_P2G_ZERO_FIRST = 1u, PIC_P2G_DETERMINISTIC = 2u };

/* Uniform local mesh (Picasso_UniformMesh.hpp:84-190): `low` is the physical
 * position of local node (0,0,0), ghost layers included; `h` the isotropic
 * cell size.  Entity lattices sit at low + h/2 in their cell-centred
 * dimensions. */
typedef struct pic_mesh {
    double low[3];
    double h;
} pic_mesh;

/* A window of a grid field: local indices [lo, lo+n) in each dimension.
 * Element (i,j,k,c) lives at
 *   data[(i-lo[0])*stride[0] + (j-lo[1])*stride[1] + (k-lo[2])*stride[2]
 *        + c*cstride]
 * (elements of the call's dtype).  Stencils must stay inside the window. */
typedef struct pic_grid {
    void* data;
    int32_t lo[3];
    int32_t n[3];
    int64_t stride[3];
    int64_t cstride;
} pic_grid;

/* A particle field.  Element (p, c) lives at
 *   data[(p >> vlen_log2) * tile_stride + (p & ((1 << vlen_log2) - 1))
 *        + c * cstride]
 * SoA: vlen_log2 = 0, tile_stride = 1, cstride = np.  AoS with q components:
 * vlen_log2 = 0, tile_stride = q, cstride = 1.  Cabana AoSoA with vector
 * length V and a struct of q scalars (Picasso_ParticleList.hpp:90-138):
 * vlen_log2 = log2 V, tile_stride = V*q, cstride = V (component offsets
 * folded into data).  Matrix fields are row-major: APIC c_p component
 * r*3 + d is c_p(r, d); PolyPIC modes component r*ncomp + d. */
typedef struct pic_pfield {
    void* data;
    int32_t vlen_log2;
    int64_t tile_stride;
    int64_t cstride;
} pic_pfield;

/* Binning geometry.  Bins are the cells of a lattice aligned with the mesh:
 * bin index c = floor((x - low)/h + dual_d/2) in dimension d, so dual_d = 0
 * bins by mesh cell and dual_d = 1 by node-centred (dual) cell -- the
 * stencil base of quadratic splines on nodes.  dual = 0: no dimension is
 * dual; dual = 1: all three are; dual = PIC_DUAL_DIMS | mask: dimension d
 * is dual when bit d of mask is set (staggered Face/Edge lattices, whose
 * knot bases are node-like in some dimensions and cell-like in others).
 * Bins [cell_lo, cell_lo+ncell) are grouped in bricks of brick[3] bins;
 * key = brick * brick_volume + bin_in_brick, both linearised i-major (k
 * fastest). */
#define PIC_DUAL_DIMS 0x10
typedef struct pic_binning {
    int32_t cell_lo[3];
    int32_t ncell[3];
    int32_t brick[3];
    int32_t dual;
} pic_binning;

/* Particles binned by pic_bin_sort: perm[s] is the particle in sorted slot
 * s, keys[s] its key; brick b holds slots [bin_start[b], bin_start[b+1]).
 * active (optional, may be NULL) lists the non-empty bricks: active[0] is
 * their number, active[1..] their indices in increasing order. */
typedef struct pic_bins {
    const uint32_t* perm;
    const uint32_t* keys;
    const uint32_t* bin_start;
    const uint32_t* active;
    pic_binning geom;
} pic_bins;

typedef struct pic_context pic_context;

/* Context: device, stream, device-side error word, scratch memory. */
int pic_context_create(pic_context** ctx, int device);
int pic_context_destroy(pic_context* ctx);
/* hipStream_t passed as void*; NULL selects the null stream. */
int pic_set_stream(pic_context* ctx, void* stream);
/* Wait for the stream; return the latched device errors and clear them:
 * PIC_ERR_BOUNDS when a stencil or particle left its window or bins,
 * otherwise PIC_ERR_CAPACITY when a deterministic p2g could not order its
 * stale particles (see PIC_P2G_DETERMINISTIC). */
int pic_synchronize(pic_context* ctx);
const char* pic_status_string(int status);
int pic_api_version(void);

/* APIC::p2g (Picasso_APIC.hpp:55-140 orders 2/3 Node/Cell; :148-232 orders
 * 2/3 Face/Edge; :240-319 order 1 Node/Cell; :327-402 order 1 Face/Edge).
 * Reference signature p2g(m_p, c_p, m_i, mu_i, sd): x and the mesh stand in
 * for the spline data sd (createSpline, Picasso_ParticleInterpolation.hpp:
 * 88-105).  c_p has 12 components (4x3).  mu_i has 3 components for
 * Node/Cell, 1 for Face/Edge.  bins may be NULL.  flags: PIC_P2G_ZERO_FIRST,
 * PIC_P2G_DETERMINISTIC. */
int pic_apic_p2g(pic_context* ctx, int order, int entity, int dtype,
                 const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                 const pic_pfield* m_p, const pic_pfield* c_p,
                 const pic_grid* m_i, const pic_grid* mu_i,
                 const pic_bins* bins, uint32_t flags);

/* APIC::g2p (Picasso_APIC.hpp:407-472 Node/Cell; :477-544 Face/Edge).
 * Reference signature g2p(u_i, c_p, sd).  Overwrites all 12 components. */
int pic_apic_g2p(pic_context* ctx, int order, int entity, int dtype,
                 const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                 const pic_grid* u_i, const pic_pfield* c_p,
                 const pic_bins* bins);

/* APIC::g2p of the grid velocity u_i = mu_i / m_i (0 where m_i == 0),
 * computed on the fly from grid mass and momentum: grid normalisation
 * (doc/basic/explict_nearly_incompressible_flow.md:18-23) fused into
 * APIC::g2p (Picasso_APIC.hpp:407-544), without materialising u_i.  The
 * division runs in FP64 for both dtypes.  m_i and mu_i share one window. */
int pic_apic_g2p_normalized(pic_context* ctx, int order, int entity,
                            int dtype, const pic_mesh* mesh, int64_t np,
                            const pic_pfield* x, const pic_grid* m_i,
                            const pic_grid* mu_i, const pic_pfield* c_p,
                            const pic_bins* bins);

/* APIC::g2p (Picasso_APIC.hpp:407-472) fused with the velocity gradient
 * L = sum_n u_n (x) grad w_n (G2P::gradient, Picasso_ParticleInterpolation.hpp:
 * 165-181; L(a,b) = du_a/dx_b) and the deformation-gradient update
 * F_p <- (I + dt L) F_p, det(F_p), F_p^{-1} in FP64 (Picasso_BatchedLinearAlgebra.hpp:
 * 3806-3834 product, :4335-4350 det, :4624-4685 inverse), in one pass over
 * the particles.  Node/Cell entities.  m_i NULL: u_i is the grid velocity;
 * otherwise u_i is the grid momentum and the velocity mu_i / m_i is formed
 * on the fly as in pic_apic_g2p_normalized.  F, det, Finv are double (9, 1,
 * 9 components); F is updated in place. */
int pic_apic_g2p_update_F(pic_context* ctx, int order, int entity, int dtype,
                          const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                          const pic_grid* m_i, const pic_grid* u_i,
                          const pic_pfield* c_p, double dt, const pic_pfield* F,
                          const pic_pfield* det, const pic_pfield* Finv,
                          const pic_bins* bins);

/* One APIC transfer step in one call (no reference function: a Picasso
 * step is the application's own sequence of GridOperator apply calls whose
 * particle loops call APIC::p2g and APIC::g2p, Picasso_GridOperator.hpp:
 * 459-496, 520-544).  Same results as
 *   pic_bin_sort(x by geom into keys, perm, bin_start, active);
 *   pic_apic_p2g(..., m_i, mu_i, those bins, PIC_P2G_ZERO_FIRST);
 *   pic_apic_g2p_normalized(..., m_i, mu_i, c_p, those bins);
 * with fewer launches when geom has 4x4x4 bricks on the spline's knot
 * lattice (the p2g brick kernels): the grid windows (when each is one dense
 * range) are cleared by extra workgroups of the sort's key pass, beside the
 * keys, and the sort's scan lists the brick walk once for both transfers.
 * Other binnings run the three calls.  Arguments as in those calls; active
 * may be NULL. */
int pic_apic_step(pic_context* ctx, int order, int entity, int dtype,
                  const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                  const pic_pfield* m_p, const pic_pfield* c_p,
                  const pic_grid* m_i, const pic_grid* mu_i,
                  const pic_binning* geom, uint32_t* keys, uint32_t* perm,
                  uint32_t* bin_start, uint32_t* active);

/* Number of PolyPIC modes per component for a spline order: 8 (linear, the
 * reference's), 27 (orders 2 and 3, the {1,y,y^2}^3 basis). */
int pic_polypic_nmode(int order);

/* PolyPIC::p2g (Picasso_PolyPIC.hpp:39-130 Node/Cell; :138-229 Face/Edge).
 * Reference signature p2g(m_p, u_p, c_p, mu_i, m_i, dt, sd) -- field before
 * mass.  u_p has nmode*3 components, c_p nmode*ncomp; mu_i has ncomp
 * components for Node/Cell and 1 for Face/Edge (component entity dim when
 * ncomp == 3, else 0). */
int pic_polypic_p2g(pic_context* ctx, int order, int entity, int dtype,
                    const pic_mesh* mesh, int64_t np, int ncomp,
                    const pic_pfield* x, const pic_pfield* m_p,
                    const pic_pfield* u_p, const pic_pfield* c_p,
                    const pic_grid* mu_i, const pic_grid* m_i, double dt,
                    const pic_bins* bins, uint32_t flags);

/* PolyPIC::g2p (Picasso_PolyPIC.hpp:235-293 Node/Cell; :299-359 Face/Edge). */
int pic_polypic_g2p(pic_context* ctx, int order, int entity, int dtype,
                    const pic_mesh* mesh, int64_t np, int ncomp,
                    const pic_pfield* x, const pic_grid* u_i,
                    const pic_pfield* c_p, const pic_bins* bins);

/* Grid halo on one GPU: the gather and scatter-sum halo steps GridOperator
 * runs around each apply (Picasso_GridOperator.hpp:203-215, 465-491;
 * Cabana::Grid::Halo, not in the reference tree), for periodic dimensions.
 * The window of g holds the owned nodes [own_lo, own_lo + own_n) and ghost
 * layers; in a dimension with periodic[d] != 0 node i is the image of
 * own_lo + ((i - own_lo) mod own_n), the ghost layers together no wider
 * than own_n.
 *  - PIC_HALO_SCATTER_SUM: every ghost adds its ncomp values into its
 *    owned image (after a p2g, the ghosts' partial sums).  Ghost values
 *    are unspecified afterwards (a gather restores them).
 *  - PIC_HALO_GATHER: every ghost takes its owned image's values (before a
 *    g2p).
 * Non-periodic dimensions exchange nothing (no neighbour on one GPU). */
enum { PIC_HALO_GATHER = 0, PIC_HALO_SCATTER_SUM = 1 };
int pic_grid_halo(pic_context* ctx, int op, int dtype, const pic_grid* g, int ncomp,
                  const int32_t own_lo[3], const int32_t own_n[3], const int32_t periodic[3]);

/* GridOperator's apply for a particle operation on one GPU
 * (Picasso_GridOperator.hpp:459-496 applyImpl).  A C ABI cannot carry the
 * reference's functor, so the operation is a host callback that issues the
 * transfer calls (APIC / PolyPIC / G2P / P2G) on the context's stream.  In
 * order: every gather field's halo is gathered (pic_grid_halo GATHER over
 * its periodic dimensions, :465-466); every scatter field is zeroed over its
 * window (:532-533); body(ctx, user) runs (:486-487); every scatter field's
 * halo is scatter-summed (:489-491).  A field without periodic dimensions
 * exchanges nothing (no neighbour on one GPU).  A non-zero return of body
 * stops the apply and is returned. */
typedef struct pic_field_dep {
    pic_grid grid;        /* the field's window, ghost layers included */
    int32_t ncomp;
    int32_t own_lo[3];    /* its owned entities, as for pic_grid_halo */
    int32_t own_n[3];
    int32_t periodic[3];
} pic_field_dep;
typedef int (*pic_apply_fn)(pic_context* ctx, void* user);
int pic_grid_operator_apply(pic_context* ctx, int dtype, int ngather,
                            const pic_field_dep* gather, int nscatter,
                            const pic_field_dep* scatter, pic_apply_fn body,
                            void* user);

/* Grid velocity u_i = mu_i / m_i over the window of m_i (no reference
 * function; doc/basic/explict_nearly_incompressible_flow.md:18-23 without
 * forces).  Nodes with m_i == 0 get u_i = 0.  For PIC_F32, nodes with
 * m_i < rel_mass_eps * max(m_i) divide in FP64. */
int pic_grid_normalize(pic_context* ctx, int dtype, int ncomp,
                       const pic_grid* m_i, const pic_grid* mu_i,
                       const pic_grid* u_i, double rel_mass_eps);

/* Velocity gradient at the particles, gradu(i,j) = du_i/dx_j (component
 * i*3+j), from a 3-component Node/Cell grid field (G2P::gradient,
 * Picasso_ParticleInterpolation.hpp:165-181). */
int pic_g2p_gradu(pic_context* ctx, int order, int entity, int dtype,
                  const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                  const pic_grid* u_i, const pic_pfield* gradu,
                  const pic_bins* bins);

/* Generic spline transfers: batch forms of Picasso::G2P::value / gradient /
 * divergence (Picasso_ParticleInterpolation.hpp:110-189) and
 * Picasso::P2G::value / gradient / divergence (:197-279), on any entity
 * lattice and spline order.  ncomp counts the input's components; grid
 * fields keep their components at stride cstride, particle fields at
 * component stride.  With w_n the node weight and gw_n its gradient:
 *   G2P VALUE       out(p, c)      = sum_n w_n f(n, c)            ncomp 1..3
 *   G2P GRADIENT    out(p, 3c + d) = sum_n gw_n[d] f(n, c)        ncomp 1 or 3
 *                   (the vector case is du_c/dx_d, as pic_g2p_gradu)
 *   G2P DIVERGENCE  out(p)         = sum_n sum_d gw_n[d] f(n, d)  ncomp 3
 *   P2G VALUE       f(n, c)       += w_n v(p, c)                  ncomp 1..3
 *   P2G GRADIENT    f(n, d)       += gw_n[d] v(p)                 ncomp 1
 *   P2G DIVERGENCE  f(n)          += sum_d gw_n[d] v(p, d)        ncomp 3
 *                   f(n, i)       += sum_j gw_n[j] v(p, 3i + j)   ncomp 9
 * P2G accumulates like a ScatterView (PIC_P2G_ZERO_FIRST zeroes the field
 * first).  bins (optional, from pic_bin_sort) only sets the visiting
 * order. */
enum { PIC_INTERP_VALUE = 0, PIC_INTERP_GRADIENT = 1, PIC_INTERP_DIVERGENCE = 2 };
int pic_g2p_interp(pic_context* ctx, int op, int order, int entity, int dtype,
                   const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                   const pic_grid* field, int ncomp, const pic_pfield* out,
                   const pic_bins* bins);
int pic_p2g_interp(pic_context* ctx, int op, int order, int entity, int dtype,
                   const pic_mesh* mesh, int64_t np, const pic_pfield* x,
                   const pic_pfield* value, int ncomp, const pic_grid* field,
                   uint32_t flags, const pic_bins* bins);

/* F_p <- (I + dt*gradu) F_p, det(F_p) and F_p^{-1} in FP64
 * (Picasso_BatchedLinearAlgebra.hpp:3806-3834 product, :4335-4350 det,
 * :4624-4685 inverse).  gradu is of gradu_dtype; F, det, Finv are double. */
int pic_update_F(pic_context* ctx, int64_t np, int gradu_dtype,
                 const pic_pfield* gradu, double dt, const pic_pfield* F,
                 const pic_pfield* det, const pic_pfield* Finv);

/* Number of bricks of a binning (the bin_start array needs one more entry);
 * a negative pic_status on invalid sizes. */
int64_t pic_bin_count(const pic_binning* geom);

/* Stable counting sort of particles by bin key (no reference function;
 * Picasso has no particle sort).  Writes perm[np] (sorted slot -> particle),
 * keys[np] (key of each sorted slot, non-decreasing) and
 * bin_start[nbricks+1]; equal keys keep their input order.  When active is
 * not NULL (nbricks+1 entries) it receives the list of non-empty bricks
 * described at pic_bins.  A particle outside the bins is reported by
 * pic_synchronize (outputs then undefined). */
int pic_bin_sort(pic_context* ctx, int dtype, const pic_mesh* mesh,
                 int64_t np, const pic_pfield* x, const pic_binning* geom,
                 uint32_t* keys, uint32_t* perm, uint32_t* bin_start,
                 uint32_t* active);

/* pic_bin_sort without any device allocation (no reference counterpart:
 * for callers that cannot allocate -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, results
 * and errors as pic_bin_sort, with the sort's scratch in one array of
 * PIC_NOALLOC_SORT_BYTES that the library keeps on every GPU it runs on until
 * the process ends, instead of the context's growable scratch.  Returns
 * PIC_ERR_CAPACITY when the sort needs more than that: the scratch takes
 * 4 * (3 np + 2 (nbricks + 1) + 2 + 2 ceil((nbricks + 1) / 2048)) bytes, plus
 * 4 (nbricks + 1) when active is NULL; np <= 4,194,304 with nbricks <
 * 1,048,576 always fits.  Every call on one GPU shares that GPU's array
 * (whatever the context): calls that overlap in time on one GPU race. */
#define PIC_NOALLOC_SORT_BYTES (64u << 20)
int pic_bin_sort_noalloc(pic_context* ctx, int dtype, const pic_mesh* mesh,
                         int64_t np, const pic_pfield* x,
                         const pic_binning* geom, uint32_t* keys,
                         uint32_t* perm, uint32_t* bin_start,
                         uint32_t* active);

/* Gather particle fields into bin order (no reference function; Picasso
 * never reorders its particles): for each of nfields fields, dst(s, c) =
 * src(perm[s], c) for s < np and c < ncomp[field].  src and dst must not
 * overlap.  With perm from pic_bin_sort this stores the particles bin by bin,
 * after which the brick kernels read them with unit-stride loads (a PIC code
 * re-sorting its particle arrays); the sort's perm then becomes nearly the
 * identity on later steps. */
int pic_permute(pic_context* ctx, int dtype, int64_t np, const uint32_t* perm,
                int nfields, const pic_pfield* src, const pic_pfield* dst,
                const int* ncomp);

/* Particle initialisation, Picasso's initializeParticles
 * (Picasso_ParticleInit.hpp:121-248 InitRandom, :279-400 InitUniform)
 * without its creation functor.  Fills every cell of [cell_lo,
 * cell_lo + ncell) (local cell indices; node i sits at mesh->low + i h):
 *  - PIC_INIT_UNIFORM: ppc^3 particles per cell on the lattice
 *    low + (0.5 + i_p) spacing, spacing = (high - low) / ppc per dimension
 *    (:345-356); ppc is particles_per_cell_dim;
 *  - PIC_INIT_RANDOM: ppc particles per cell, each coordinate drawn
 *    uniformly in [low, high) (:193-197) from a counter-based generator of
 *    (seed, particle, dimension) -- the same positions for the same seed,
 *    whatever the device; the reference's Kokkos XorShift pool is only
 *    specified by its distribution.
 * Particle p = cell_id * n + local with cell_id = i + ncell[0] (j + ncell[1]
 * k) (i fastest, :170-173) and, uniform, local = ip + ppc (jp + ppc kp)
 * (:339-341).  x receives 3 components; volume (may be NULL) 1, the cell
 * volume / n (:200).  Returns the number of particles written (<= capacity)
 * or a negative status (PIC_ERR_CAPACITY when capacity is too small).  The
 * creation functor's choice is then a keep mask: pic_select + pic_permute
 * compact the fields (filterEmpties, :38-91). */
enum { PIC_INIT_UNIFORM = 0, PIC_INIT_RANDOM = 1 };
int64_t pic_init_particles(pic_context* ctx, int mode, int dtype, const pic_mesh* mesh,
                           const int32_t cell_lo[3], const int32_t ncell[3], int ppc,
                           uint64_t seed, int64_t capacity, const pic_pfield* x,
                           const pic_pfield* volume);

/* Stable selection (the compaction of Picasso's filterEmpties,
 * Picasso_ParticleInit.hpp:38-91, which keeps the created particles but
 * moves tail particles into the holes; here the order is kept): kept[0]
 * receives the number of i < np with keep[i] != 0 and kept[1..] those i in
 * increasing order.  kept holds np + 1 words.  pic_permute with
 * perm = kept + 1 then gathers the particle fields. */
int pic_select(pic_context* ctx, int64_t np, const uint8_t* keep, uint32_t* kept);

/* Step graphs (no reference counterpart: Kokkos launches each kernel).
 * Between pic_graph_begin and pic_graph_end the context's calls are
 * captured into a HIP graph instead of running; pic_graph_launch replays
 * the whole sequence on the context's stream with one launch.  The context
 * must be bound to a stream other than the null stream.  A captured call
 * may not need more scratch than the context already holds (run the same
 * calls once before capturing); such a call returns PIC_ERR_ALLOC.  No
 * kernel timing is recorded inside a capture; pic_synchronize and a change
 * of stream are refused until pic_graph_end. */
typedef struct pic_graph pic_graph;
int pic_graph_begin(pic_context* ctx);
int pic_graph_end(pic_context* ctx, pic_graph** graph);
int pic_graph_launch(pic_context* ctx, const pic_graph* graph);
int pic_graph_destroy(pic_graph* graph);

/* Batches of independent instances (no reference counterpart: Picasso runs
 * one simulation per GridOperator, whose apply launches one Kokkos kernel
 * per operator, Picasso_GridOperator.hpp:632-641).  An instance is one APIC
 * simulation on Node splines with its own local mesh, its particles
 * [first, first + np) of the batch's particle fields (instance k + 1 starts
 * where instance k ends), its binning (4x4x4 bricks on the spline's knot
 * lattice: dual 1 for quadratic splines, 0 for linear and cubic) and its
 * grid windows: mass m_i, and momentum mu_i with 3 components over the same
 * window, each window one dense range of memory.  pic_batch_create checks
 * and uploads the instance table and allocates everything the batch's calls
 * use (bins, walk lists, scratch), so those calls allocate nothing and a
 * whole step can be captured in one step graph.  Each call is one launch
 * per pass over the bricks of every instance.  Errors detected on the
 * device are reported by pic_synchronize as for the other calls. */
typedef struct pic_instance {
    pic_mesh mesh;
    int64_t first;
    int64_t np;
    pic_binning geom;
    pic_grid m_i;
    pic_grid mu_i;
} pic_instance;

typedef struct pic_batch pic_batch;
int pic_batch_create(pic_context* ctx, int order, int dtype, int ninst,
                     const pic_instance* inst, pic_batch** batch);
int pic_batch_destroy(pic_batch* batch);
/* pic_bin_sort of every instance's particles by its own bins, in one pass
 * over all particles.  x holds the particles of every instance. */
int pic_batch_bin_sort(pic_context* ctx, pic_batch* batch, const pic_pfield* x);
/* APIC::p2g (Picasso_APIC.hpp:55-140, 240-319) of every instance into its
 * own grids through the bins of the last pic_batch_bin_sort;
 * PIC_P2G_ZERO_FIRST zeroes every instance's windows first. */
int pic_batch_apic_p2g(pic_context* ctx, pic_batch* batch, const pic_pfield* x,
                       const pic_pfield* m_p, const pic_pfield* c_p, uint32_t flags);
/* APIC::g2p (Picasso_APIC.hpp:407-472) of every instance from its grid
 * velocity mu_i / m_i formed on the fly, as pic_apic_g2p_normalized. */
int pic_batch_apic_g2p_normalized(pic_context* ctx, pic_batch* batch, const pic_pfield* x,
                                  const pic_pfield* c_p);
/* One APIC step of every instance: pic_batch_bin_sort, pic_batch_apic_p2g,
 * pic_batch_apic_g2p_normalized. */
int pic_batch_apic_step(pic_context* ctx, pic_batch* batch, const pic_pfield* x,
                        const pic_pfield* m_p, const pic_pfield* c_p, uint32_t flags);

/* Tuning options of a context (no reference counterpart; picasso_amd/tune.py
 * picks them on a live workload).  PIC_OPT_G2P_GROUP: work-items per
 * workgroup of the collocated APIC g2p brick kernels without the F update
 * (pic_apic_g2p, pic_apic_g2p_normalized, pic_apic_step and the batches'
 * g2p), 128 (four particles per work-item at a time) or 256 (three); 0
 * restores the default (256).
 * Results do not depend on it. */
enum { PIC_OPT_G2P_GROUP = 1 };
int pic_set_option(pic_context* ctx, int option, int value);

/* Kernel timing: when enabled, every launch is bracketed by HIP events on
 * the context's stream and pic_kernel_time() returns the accumulated
 * milliseconds and launch count of the named kernel family
 * ("apic_p2g", "apic_g2p", "normalize", "bin_sort", ...). */
int pic_set_timing(pic_context* ctx, int enable);
int pic_kernel_time(pic_context* ctx, const char* name, double* ms,
                    int64_t* launches);
int pic_reset_timing(pic_context* ctx);

#ifdef __cplusplus
}
#endif

#endif
