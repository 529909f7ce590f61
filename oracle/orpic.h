/* orpic.h -- CPU FP64 oracle for the particle<->grid transfer path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against (tests/, the smoke check and the CPU
 * baseline leg of bench.py).  The product path never links, loads or calls it.
 *
 * It restates, in plain C with FP contraction disabled, the per-particle
 * functions of the reference (Picasso, BSD-3-Clause):
 *   src/Picasso_APIC.hpp                 APIC::inertialScaling/p2g/g2p
 *   src/Picasso_PolyPIC.hpp              PolyPIC::p2g/g2p (linear, 8 modes)
 *   src/Picasso_BatchedLinearAlgebra.hpp 3x3 det (operator!) and inverse
 *   src/Picasso_ParticleInterpolation.hpp createSpline -> Cabana evaluateSpline
 * plus the pieces the reference leaves to the caller (cell keys and a stable
 * counting sort, grid normalisation, the deformation-gradient update), each
 * specified in its function comment.
 *
 * Conventions shared by every function:
 *  - Lattice: a uniform local mesh with low corner `low` (the physical
 *    position of local node (0,0,0), ghost layers included) and isotropic cell
 *    size `h`.  Entity lattices are offset by h/2 in their cell-centred
 *    dimensions (see orc_entity_offset).
 *  - Grid arrays are dense windows: element (i,j,k,c) of a window with
 *    origin lo[3] and extent n[3] sits at ((i-lo0)*n1 + (j-lo1))*n2 + (k-lo2))
 *    * ncomp + c.  A stencil node outside the window is an error.
 *  - Particle arrays are AoS: x[p][3], m[p], APIC c[p][4][3] (row 0 velocity,
 *    rows 1..3 hold B^T: c(1+a,d) = B(d,a)), PolyPIC modes[p][nmode][ncomp].
 *  - Return 0 on success, a negative code on error.
 */
#ifndef ORPIC_H
#define ORPIC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_NODE = 0,
    ORC_CELL = 1,
    ORC_FACE_I = 2,
    ORC_FACE_J = 3,
    ORC_FACE_K = 4,
    ORC_EDGE_I = 5,
    ORC_EDGE_J = 6,
    ORC_EDGE_K = 7
};

enum {
    ORC_OK = 0,
    ORC_ERR_ARG = -1,
    ORC_ERR_BOUNDS = -2,
    ORC_ERR_ALLOC = -3
};

int orc_version(void);

/* Physical offset (0 or h/2) of the entity lattice in dimension d. */
double orc_entity_offset(int entity, int d, double h);

/* Spline data for one particle: node indices s, weights w, physical weight
 * gradients g, physical second derivatives g2 and physical distances
 * d = x_node - x_p, each [3][4] (row = dimension, first nk entries used).
 * Returns the number of knots per dimension (order+1) or a negative code. */
int orc_spline(int order, int entity, const double low[3], double h,
               const double x[3], int32_t* s, double* w, double* g,
               double* g2, double* d);

/* APIC p2g (Picasso_APIC.hpp:55-140 orders 2/3 collocated, :148-232
 * staggered, :240-319 order 1 collocated, :327-402 order 1 staggered).
 * Accumulates into gm (ncomp 1) and gmu (ncomp 3 collocated, 1 staggered). */
int orc_apic_p2g(int order, int entity, const double low[3], double h,
                 const int32_t lo[3], const int32_t n[3], int64_t np,
                 const double* x, const double* m, const double* c,
                 double* gm, double* gmu);

/* APIC g2p (Picasso_APIC.hpp:407-472 collocated, :477-544 staggered).
 * Overwrites c[p][4][3]. */
int orc_apic_g2p(int order, int entity, const double low[3], double h,
                 const int32_t lo[3], const int32_t n[3], int64_t np,
                 const double* x, const double* gu, double* c);

/* Number of PolyPIC modes for a spline order (8 for linear, 27 for cubic). */
int orc_polypic_nmode(int order);

/* PolyPIC p2g (Picasso_PolyPIC.hpp:39-130 collocated, :138-229 staggered).
 * Argument order follows the reference: field (mu) before mass (m).
 * u[p][nmode][3] drives the Lagrangian backtrack, c[p][nmode][ncomp] is the
 * projected field.  Order 3 (27 modes) is this library's own extension. */
int orc_polypic_p2g(int order, int entity, const double low[3], double h,
                    const int32_t lo[3], const int32_t n[3], int64_t np,
                    const double* x, const double* m, const double* u,
                    const double* c, int ncomp, double dt, double* gmu,
                    double* gm);

/* PolyPIC g2p (Picasso_PolyPIC.hpp:235-293 collocated, :299-359 staggered). */
int orc_polypic_g2p(int order, int entity, const double low[3], double h,
                    const int32_t lo[3], const int32_t n[3], int64_t np,
                    const double* x, const double* gu, int ncomp, double* c);

/* Generic G2P / P2G transfers (value, gradient, divergence); see orpic.c. */
enum { ORC_INTERP_VALUE = 0, ORC_INTERP_GRADIENT = 1, ORC_INTERP_DIVERGENCE = 2 };
int orc_g2p_interp(int op, int order, int entity, const double low[3], double h,
                   const int32_t lo[3], const int32_t n[3], int64_t np, const double* x,
                   const double* f, int ncomp, double* out);
int orc_p2g_interp(int op, int order, int entity, const double low[3], double h,
                   const int32_t lo[3], const int32_t n[3], int64_t np, const double* x,
                   const double* v, int ncomp, double* f);

/* Velocity gradient at the particle, gradu[p][i][j] = du_i/dx_j, from a
 * collocated 3-component grid field (G2P::gradient,
 * Picasso_ParticleInterpolation.hpp:165-181). */
int orc_g2p_gradu(int order, int entity, const double low[3], double h,
                  const int32_t lo[3], const int32_t n[3], int64_t np,
                  const double* x, const double* gu, double* gradu);

/* u = mu/m at nodes with m > 0, u = 0 elsewhere; mu,u are [nn][ncomp]. */
int orc_grid_normalize(int64_t nn, int ncomp, const double* m,
                       const double* mu, double* u);

/* Particle initialisation (Picasso_ParticleInit.hpp:121-248, 279-400)
 * without the creation functor: mode 0 = uniform lattice of ppc^3 per cell,
 * 1 = ppc random points per cell (coordinates orc_uniform01(seed, p, d));
 * x[p][3], volume[p] (may be NULL).  Returns the particle count. */
double orc_uniform01(uint64_t seed, uint64_t pid, int d);
int64_t orc_init_particles(int mode, const double low[3], double h, const int32_t cell_lo[3],
                           const int32_t ncell[3], int ppc, uint64_t seed, double* x,
                           double* volume);

/* Grid halo on one rank for periodic dimensions (GridOperator's gather /
 * scatter-sum, Picasso_GridOperator.hpp:203-215, 465-491): op 1 adds every
 * ghost into its image olo + ((i - olo) mod on), one periodic dimension
 * after the other (ghost values are unspecified afterwards); op 0 copies
 * images into ghosts.  g is [n0][n1][n2][ncomp]
 * over the window [lo, lo+n). */
int orc_grid_halo(int op, const int32_t lo[3], const int32_t n[3], int ncomp, double* g,
                  const int32_t olo[3], const int32_t on[3], const int32_t periodic[3]);

/* Cell keys.  cell = floor((x - low) * (1/h) + dual_d/2) per dimension d
 * (dual_d = 0: mesh cells, 1: node-centred dual cells; dual = 0: none,
 * 1: every dimension, 0x10 | mask: dimension d when bit d is set, as
 * pic_binning), relative to cell_lo, must lie
 * in [0, ncell).  Cells are grouped in bricks of brick[3] cells;
 * key = brick_linear * brick_volume + cell_in_brick_linear, both
 * linearisations i-major (k fastest).  brick = {1,1,1} gives the plain
 * linear cell index.  cells[p][3] receives the absolute cell indices. */
int orc_cell_keys(int64_t np, const double* x, const double low[3], double h,
                  const int32_t cell_lo[3], const int32_t ncell[3],
                  const int32_t brick[3], int dual, uint32_t* keys,
                  int32_t* cells);

/* Stable counting sort by key: perm[s] = index of the particle in sorted
 * slot s; key_start[k] = first slot of key k (nkeys+1 entries). */
int orc_bin_sort(int64_t np, const uint32_t* keys, uint32_t nkeys,
                 uint32_t* perm, uint32_t* key_start);

/* 3x3 determinant and inverse in the reference's operation order
 * (Picasso_BatchedLinearAlgebra.hpp:4335-4350, 4624-4685). */
double orc_det3(const double* a);
void orc_inverse3(const double* a, double* ainv);

/* F <- (I + dt*gradu) F, then det(F) and F^{-1} (row-major 3x3 each). */
int orc_update_F(int64_t np, const double* gradu, double dt, double* F,
                 double* det, double* Finv);

/* CPU baseline: one APIC p2g -> normalise -> g2p step on a zeroed node grid
 * window, OpenMP over particles.  mode 0: per-thread duplicated grids then a
 * reduction (Kokkos OpenMP ScatterView default); mode 1: atomic adds.
 * nthreads <= 0 uses the OpenMP default. */
int orc_apic_step_omp(int order, const double low[3], double h,
                      const int32_t lo[3], const int32_t n[3], int64_t np,
                      const double* x, const double* m, double* c,
                      double* gm, double* gmu, double* gu, int mode,
                      int nthreads);

#ifdef __cplusplus
}
#endif

#endif
