/* orpic.c -- CPU FP64 oracle for the particle<->grid transfer path.
 *
 * TEST INFRASTRUCTURE ONLY (see orpic.h): the checker the HIP implementation
 * is compared against, never part of the product path.
 *
 * Built with -ffp-contract=off so that every expression rounds as written.
 * Each function names the reference lines it restates.
 */
#include "orpic.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

int orc_version(void) { return 1; }

/* Entity lattices (Cabana::Grid conventions as used by Picasso,
 * Picasso_FieldTypes.hpp:125-168): Node is node-aligned everywhere, Cell is
 * cell-centred everywhere, Face<D> is node-aligned in D and cell-centred in the
 * other two dimensions, Edge<D> is cell-centred in D and node-aligned in the
 * other two.  The box shifts of the reference's staggered tests
 * (unit_test/tstAPIC.hpp:698-716, tstPolyPIC.hpp:351-369) line these lattices
 * up with the Mathematica grids, which pins the offsets. */
static int cell_centred(int entity, int d)
{
    switch (entity) {
    case ORC_NODE: return 0;
    case ORC_CELL: return 1;
    case ORC_FACE_I:
    case ORC_FACE_J:
    case ORC_FACE_K: return d != entity - ORC_FACE_I;
    case ORC_EDGE_I:
    case ORC_EDGE_J:
    case ORC_EDGE_K: return d == entity - ORC_EDGE_I;
    default: return 0;
    }
}

/* Component a staggered (Face/Edge) transfer works on: entity_type::dim. */
static int stag_dim(int entity)
{
    if (entity >= ORC_FACE_I && entity <= ORC_FACE_K) return entity - ORC_FACE_I;
    if (entity >= ORC_EDGE_I && entity <= ORC_EDGE_K) return entity - ORC_EDGE_I;
    return -1;
}

double orc_entity_offset(int entity, int d, double h)
{
    return cell_centred(entity, d) ? 0.5 * h : 0.0;
}

typedef struct {
    int nk;
    int32_t s[3][4];
    double w[3][4], g[3][4], g2[3][4], d[3][4];
} spl_t;

/* B-spline weights, gradients, second derivatives and distances on one
 * lattice (Picasso_ParticleInterpolation.hpp:88-105 -> Cabana evaluateSpline).
 * Definitions from the notebooks doc/apic/apic_3d_{linear,quadratic,cubic}_
 * unit_test.nb: with t = (x - o)/h the logical position,
 *   order 1: base = floor(t),       f = t - base, knots base..base+1,
 *            w = (1-f, f)
 *   order 2: base = floor(t + 1/2), f = t - base, knots base-1..base+1,
 *            w = (1/2 (1/2-f)^2, 3/4 - f^2, 1/2 (1/2+f)^2)
 *   order 3: base = floor(t),       f = t - base, knots base-1..base+2,
 *            w = ((1-f)^3/6, f^3/2 - f^2 + 2/3, (1-f)^3/2 - (1-f)^2 + 2/3,
 *                 f^3/6)
 * Gradients are d/dx = (1/h) d/dt; distances are d = x_node - x_p
 * (notebook r = nx - px). */
static int spline(int order, int entity, const double low[3], double h,
                  const double x[3], spl_t* sd)
{
    if (order < 1 || order > 3 || entity < ORC_NODE || entity > ORC_EDGE_K ||
        !(h > 0.0))
        return ORC_ERR_ARG;
    const double rdx = 1.0 / h;
    const double rdx2 = rdx * rdx;
    sd->nk = order + 1;
    for (int dim = 0; dim < 3; ++dim) {
        const double o = low[dim] + orc_entity_offset(entity, dim, h);
        const double t = (x[dim] - o) * rdx;
        double *w = sd->w[dim], *g = sd->g[dim], *g2 = sd->g2[dim];
        int first;
        double base_d;
        if (order == 1) {
            base_d = floor(t);
            const double f = t - base_d;
            first = 0;
            w[0] = 1.0 - f;
            w[1] = f;
            g[0] = -rdx;
            g[1] = rdx;
            g2[0] = 0.0;
            g2[1] = 0.0;
        } else if (order == 2) {
            base_d = floor(t + 0.5);
            const double f = t - base_d;
            const double a = 0.5 - f, b = 0.5 + f;
            first = -1;
            w[0] = 0.5 * a * a;
            w[1] = 0.75 - f * f;
            w[2] = 0.5 * b * b;
            g[0] = -a * rdx;
            g[1] = -2.0 * f * rdx;
            g[2] = b * rdx;
            g2[0] = rdx2;
            g2[1] = -2.0 * rdx2;
            g2[2] = rdx2;
        } else {
            base_d = floor(t);
            const double f = t - base_d;
            const double q = 1.0 - f;
            first = -1;
            w[0] = q * q * q / 6.0;
            w[1] = 0.5 * f * f * f - f * f + 2.0 / 3.0;
            w[2] = 0.5 * q * q * q - q * q + 2.0 / 3.0;
            w[3] = f * f * f / 6.0;
            g[0] = -0.5 * q * q * rdx;
            g[1] = (1.5 * f * f - 2.0 * f) * rdx;
            g[2] = (2.0 * q - 1.5 * q * q) * rdx;
            g[3] = 0.5 * f * f * rdx;
            g2[0] = q * rdx2;
            g2[1] = (3.0 * f - 2.0) * rdx2;
            g2[2] = (1.0 - 3.0 * f) * rdx2;
            g2[3] = f * rdx2;
        }
        const int base = (int)base_d;
        for (int n = 0; n < sd->nk; ++n) {
            sd->s[dim][n] = base + first + n;
            sd->d[dim][n] = (o + (double)(base + first + n) * h) - x[dim];
        }
    }
    return sd->nk;
}

int orc_spline(int order, int entity, const double low[3], double h,
               const double x[3], int32_t* s, double* w, double* g,
               double* g2, double* d)
{
    spl_t sd;
    int nk = spline(order, entity, low, h, x, &sd);
    if (nk < 0) return nk;
    for (int dim = 0; dim < 3; ++dim)
        for (int n = 0; n < 4; ++n) {
            int in = n < nk;
            s[dim * 4 + n] = in ? sd.s[dim][n] : 0;
            w[dim * 4 + n] = in ? sd.w[dim][n] : 0.0;
            g[dim * 4 + n] = in ? sd.g[dim][n] : 0.0;
            g2[dim * 4 + n] = in ? sd.g2[dim][n] : 0.0;
            d[dim * 4 + n] = in ? sd.d[dim][n] : 0.0;
        }
    return nk;
}

/* Dense window index of node (i,j,k), or -1 when outside the window. */
static int64_t win_index(const int32_t lo[3], const int32_t n[3], int i, int j,
                         int k)
{
    const int64_t a = (int64_t)i - lo[0], b = (int64_t)j - lo[1],
                  c = (int64_t)k - lo[2];
    if (a < 0 || b < 0 || c < 0 || a >= n[0] || b >= n[1] || c >= n[2])
        return -1;
    return (a * n[1] + b) * n[2] + c;
}

static int stencil_inside(const spl_t* sd, const int32_t lo[3],
                          const int32_t n[3])
{
    for (int d = 0; d < 3; ++d)
        if (sd->s[d][0] < lo[d] || sd->s[d][sd->nk - 1] >= lo[d] + n[d]) return 0;
    return 1;
}

/* APIC::inertialScaling (Picasso_APIC.hpp:33-47): D_p^{-1} = 4/dx^2 for
 * quadratic, 3/dx^2 for cubic splines. */
static double inertial_scaling(int order, double h)
{
    return (order == 2 ? 4.0 : 3.0) / (h * h);
}

/* Accumulation target: plain (+=) or OpenMP atomic. */
static inline void acc(double* p, double v, int atomic)
{
    if (atomic) {
#pragma omp atomic
        *p += v;
    } else {
        *p += v;
    }
}

/* One particle of APIC p2g. */
static int apic_p2g_one(int order, int entity, const double low[3], double h,
                        const int32_t lo[3], const int32_t n[3],
                        const double* xp, double mp, const double* cp,
                        double* gm, double* gmu, int atomic)
{
    spl_t sd;
    int nk = spline(order, entity, low, h, xp, &sd);
    if (nk < 0) return nk;
    if (!stencil_inside(&sd, lo, n)) return ORC_ERR_BOUNDS;
    const int sdim = stag_dim(entity);

    /* B_p is sliced from c_p and transposed (Picasso_APIC.hpp:94-101). */
    double B[3][3];
    for (int d = 0; d < 3; ++d)
        for (int a = 0; a < 3; ++a) B[d][a] = cp[(1 + a) * 3 + d];

    if (order >= 2) {
        const double Dinv = inertial_scaling(order, h);
        /* Loop order i, j, k (Picasso_APIC.hpp:110-139, 209-231). */
        for (int i = 0; i < nk; ++i)
            for (int j = 0; j < nk; ++j)
                for (int k = 0; k < nk; ++k) {
                    const double dist[3] = {sd.d[0][i], sd.d[1][j], sd.d[2][k]};
                    const double wm = sd.w[0][i] * sd.w[1][j] * sd.w[2][k] * mp;
                    const int64_t node =
                        win_index(lo, n, sd.s[0][i], sd.s[1][j], sd.s[2][k]);
                    if (sdim < 0) {
                        for (int d = 0; d < 3; ++d) {
                            /* (D_p_inv * B_p) * distance */
                            double dbd = 0.0;
                            for (int a = 0; a < 3; ++a)
                                dbd += (Dinv * B[d][a]) * dist[a];
                            acc(&gmu[node * 3 + d], wm * (cp[d] + dbd), atomic);
                        }
                    } else {
                        double dbd = 0.0;
                        for (int a = 0; a < 3; ++a)
                            dbd += (Dinv * B[sdim][a]) * dist[a];
                        acc(&gmu[node], wm * (cp[sdim] + dbd), atomic);
                    }
                    acc(&gm[node], wm, atomic);
                }
    } else {
        /* Order 1: gradient form with no inertial scaling
         * (Picasso_APIC.hpp:287-318, 377-401). */
        for (int i = 0; i < nk; ++i)
            for (int j = 0; j < nk; ++j)
                for (int k = 0; k < nk; ++k) {
                    const double wm = sd.w[0][i] * sd.w[1][j] * sd.w[2][k] * mp;
                    const double gmv[3] = {
                        sd.g[0][i] * sd.w[1][j] * sd.w[2][k] * mp,
                        sd.w[0][i] * sd.g[1][j] * sd.w[2][k] * mp,
                        sd.w[0][i] * sd.w[1][j] * sd.g[2][k] * mp};
                    const int64_t node =
                        win_index(lo, n, sd.s[0][i], sd.s[1][j], sd.s[2][k]);
                    if (sdim < 0) {
                        for (int d = 0; d < 3; ++d) {
                            double bg = 0.0;
                            for (int a = 0; a < 3; ++a) bg += B[d][a] * gmv[a];
                            acc(&gmu[node * 3 + d], wm * cp[d] + bg, atomic);
                        }
                    } else {
                        double bg = 0.0;
                        for (int a = 0; a < 3; ++a) bg += B[sdim][a] * gmv[a];
                        acc(&gmu[node], wm * cp[sdim] + bg, atomic);
                    }
                    acc(&gm[node], wm, atomic);
                }
    }
    return ORC_OK;
}

int orc_apic_p2g(int order, int entity, const double low[3], double h,
                 const int32_t lo[3], const int32_t n[3], int64_t np,
                 const double* x, const double* m, const double* c,
                 double* gm, double* gmu)
{
    for (int64_t p = 0; p < np; ++p) {
        int rc = apic_p2g_one(order, entity, low, h, lo, n, x + 3 * p, m[p],
                              c + 12 * p, gm, gmu, 0);
        if (rc < 0) return rc;
    }
    return ORC_OK;
}

/* One particle of APIC g2p (Picasso_APIC.hpp:407-472, 477-544). */
static int apic_g2p_one(int order, int entity, const double low[3], double h,
                        const int32_t lo[3], const int32_t n[3],
                        const double* xp, const double* gu, double* cp)
{
    spl_t sd;
    int nk = spline(order, entity, low, h, xp, &sd);
    if (nk < 0) return nk;
    if (!stencil_inside(&sd, lo, n)) return ORC_ERR_BOUNDS;
    const int sdim = stag_dim(entity);
    double up[3] = {0.0, 0.0, 0.0};
    double B[3][3] = {{0.0}};
    for (int i = 0; i < nk; ++i)
        for (int j = 0; j < nk; ++j)
            for (int k = 0; k < nk; ++k) {
                const double wip = sd.w[0][i] * sd.w[1][j] * sd.w[2][k];
                const int64_t node =
                    win_index(lo, n, sd.s[0][i], sd.s[1][j], sd.s[2][k]);
                const double dist[3] = {sd.d[0][i], sd.d[1][j], sd.d[2][k]};
                if (sdim < 0) {
                    for (int d = 0; d < 3; ++d) {
                        const double ue = gu[node * 3 + d];
                        up[d] = up[d] + wip * ue;
                        for (int a = 0; a < 3; ++a) B[d][a] += (wip * ue) * dist[a];
                    }
                } else {
                    const double ue = gu[node];
                    up[sdim] += wip * ue;
                    for (int a = 0; a < 3; ++a) B[sdim][a] += (wip * ue) * dist[a];
                }
            }
    /* c_p is rebuilt from u_p and B_p^T; a staggered g2p leaves zeros in
     * the other components (Picasso_APIC.hpp:505-543). */
    for (int d = 0; d < 3; ++d) {
        cp[d] = up[d];
        for (int a = 0; a < 3; ++a) cp[(1 + a) * 3 + d] = B[d][a];
    }
    return ORC_OK;
}

int orc_apic_g2p(int order, int entity, const double low[3], double h,
                 const int32_t lo[3], const int32_t n[3], int64_t np,
                 const double* x, const double* gu, double* c)
{
    for (int64_t p = 0; p < np; ++p) {
        int rc = apic_g2p_one(order, entity, low, h, lo, n, x + 3 * p, gu,
                              c + 12 * p);
        if (rc < 0) return rc;
    }
    return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* 3x3 linear algebra. */

/* operator! 3x3 (Picasso_BatchedLinearAlgebra.hpp:4335-4350), same term
 * order. */
double orc_det3(const double* a)
{
#define A(i, j) a[(i) * 3 + (j)]
    return A(0, 0) * A(1, 1) * A(2, 2) + A(0, 1) * A(1, 2) * A(2, 0) +
           A(0, 2) * A(1, 0) * A(2, 1) - A(0, 2) * A(1, 1) * A(2, 0) -
           A(0, 1) * A(1, 0) * A(2, 2) - A(0, 0) * A(1, 2) * A(2, 1);
}

/* inverse 3x3 (Picasso_BatchedLinearAlgebra.hpp:4624-4685): adjugate times
 * a_det_inv = 1/det, same cofactor expressions. */
void orc_inverse3(const double* a, double* r)
{
    const double di = 1.0 / orc_det3(a);
    r[0] = (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) * di;
    r[1] = (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) * di;
    r[2] = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) * di;
    r[3] = (A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2)) * di;
    r[4] = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) * di;
    r[5] = (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) * di;
    r[6] = (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) * di;
    r[7] = (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) * di;
    r[8] = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) * di;
#undef A
}

/* ---------------------------------------------------------------------- */
/* PolyPIC. */

/* Mode exponents (a_x, a_y, a_z).  The first eight are the reference's
 * linear modes in its order (Picasso_PolyPIC.hpp:108-116, 269-284):
 * 1, y0, y1, y2, y0y1, y0y2, y1y2, y0y1y2.  The 27-mode set {1, y, y^2}^3
 * appends the 19 modes holding a square, in lexicographic order. */
static const int MODE_EXP[27][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1},
    {0, 1, 1}, {1, 1, 1}, {0, 0, 2}, {0, 1, 2}, {0, 2, 0}, {0, 2, 1},
    {0, 2, 2}, {1, 0, 2}, {1, 1, 2}, {1, 2, 0}, {1, 2, 1}, {1, 2, 2},
    {2, 0, 0}, {2, 0, 1}, {2, 0, 2}, {2, 1, 0}, {2, 1, 1}, {2, 1, 2},
    {2, 2, 0}, {2, 2, 1}, {2, 2, 2}};

int orc_polypic_nmode(int order) { return order == 1 ? 8 : 27; }

static double ipow(double y, int e) { return e == 0 ? 1.0 : (e == 1 ? y : y * y); }

/* PolyPIC p2g (Picasso_PolyPIC.hpp:39-130, 138-229).  Linear (8 modes) is
 * the reference's; the 27-mode basis on other orders uses the same
 * construction with the basis extended to {1, y, y^2}^3. */
static int polypic_p2g_one(int order, int entity, const double low[3],
                           double h, const int32_t lo[3], const int32_t n[3],
                           const double* xp, double mp, const double* up,
                           const double* cp, int ncomp, double dt, double* gmu,
                           double* gm)
{
    const int nmode = orc_polypic_nmode(order);
    spl_t sd;
    int nk = spline(order, entity, low, h, xp, &sd);
    if (nk < 0) return nk;
    if (!stencil_inside(&sd, lo, n)) return ORC_ERR_BOUNDS;
    const int sdim = stag_dim(entity);
    const int dim = sdim < 0 ? -1 : (ncomp == 3 ? sdim : 0);

    /* Affine material motion operator (Picasso_PolyPIC.hpp:81-84). */
#define U(r, d) up[(r) * 3 + (d)]
    const double am[9] = {1.0 + dt * U(1, 0), dt * U(1, 1), dt * U(1, 2),
                          dt * U(2, 0), 1.0 + dt * U(2, 1), dt * U(2, 2),
                          dt * U(3, 0), dt * U(3, 1), 1.0 + dt * U(3, 2)};
#undef U
    double ai[9];
    orc_inverse3(am, ai);

    for (int i = 0; i < nk; ++i)
        for (int j = 0; j < nk; ++j)
            for (int k = 0; k < nk; ++k) {
                const double wm = mp * sd.w[0][i] * sd.w[1][j] * sd.w[2][k];
                const double dist[3] = {sd.d[0][i], sd.d[1][j], sd.d[2][k]};
                double y[3];
                for (int a = 0; a < 3; ++a)
                    y[a] = ai[a * 3 + 0] * dist[0] + ai[a * 3 + 1] * dist[1] +
                           ai[a * 3 + 2] * dist[2];
                double basis[27];
                for (int r = 0; r < nmode; ++r)
                    basis[r] = ipow(y[0], MODE_EXP[r][0]) *
                               ipow(y[1], MODE_EXP[r][1]) *
                               ipow(y[2], MODE_EXP[r][2]);
                /* The reference spells the linear basis out term by term
                 * (Picasso_PolyPIC.hpp:109-116); these products round the
                 * same way. */
                const int64_t node =
                    win_index(lo, n, sd.s[0][i], sd.s[1][j], sd.s[2][k]);
                if (dim < 0) {
                    for (int d = 0; d < ncomp; ++d) {
                        /* wm * ~c_p.column(d) * basis */
                        double v = 0.0;
                        for (int r = 0; r < nmode; ++r)
                            v += (wm * cp[r * ncomp + d]) * basis[r];
                        gmu[node * ncomp + d] += v;
                    }
                } else {
                    double v = 0.0;
                    for (int r = 0; r < nmode; ++r)
                        v += (wm * cp[r * ncomp + dim]) * basis[r];
                    gmu[node] += v;
                }
                gm[node] += wm;
            }
    return ORC_OK;
}

int orc_polypic_p2g(int order, int entity, const double low[3], double h,
                    const int32_t lo[3], const int32_t n[3], int64_t np,
                    const double* x, const double* m, const double* u,
                    const double* c, int ncomp, double dt, double* gmu,
                    double* gm)
{
    if (ncomp < 1 || ncomp > 3) return ORC_ERR_ARG;
    const int nmode = orc_polypic_nmode(order);
    for (int64_t p = 0; p < np; ++p) {
        int rc = polypic_p2g_one(order, entity, low, h, lo, n, x + 3 * p, m[p],
                                 u + (int64_t)nmode * 3 * p,
                                 c + (int64_t)nmode * ncomp * p, ncomp, dt, gmu,
                                 gm);
        if (rc < 0) return rc;
    }
    return ORC_OK;
}

/* Per-dimension derivative factor of order e for g2p mode coefficients:
 * value, gradient, and half the second derivative (Taylor coefficient). */
static double dfac(const spl_t* sd, int dim, int n, int e)
{
    return e == 0 ? sd->w[dim][n]
                  : (e == 1 ? sd->g[dim][n] : 0.5 * sd->g2[dim][n]);
}

/* PolyPIC g2p (Picasso_PolyPIC.hpp:235-293, 299-359). */
static int polypic_g2p_one(int order, int entity, const double low[3],
                           double h, const int32_t lo[3], const int32_t n[3],
                           const double* xp, const double* gu, int ncomp,
                           double* cp)
{
    const int nmode = orc_polypic_nmode(order);
    spl_t sd;
    int nk = spline(order, entity, low, h, xp, &sd);
    if (nk < 0) return nk;
    if (!stencil_inside(&sd, lo, n)) return ORC_ERR_BOUNDS;
    const int sdim = stag_dim(entity);
    const int dim = sdim < 0 ? -1 : (ncomp == 3 ? sdim : 0);

    if (dim < 0) {
        for (int r = 0; r < nmode * ncomp; ++r) cp[r] = 0.0;
    } else {
        for (int r = 0; r < nmode; ++r) cp[r * ncomp + dim] = 0.0;
    }
    for (int i = 0; i < nk; ++i)
        for (int j = 0; j < nk; ++j)
            for (int k = 0; k < nk; ++k) {
                double coeff[27];
                for (int r = 0; r < nmode; ++r)
                    coeff[r] = dfac(&sd, 0, i, MODE_EXP[r][0]) *
                               dfac(&sd, 1, j, MODE_EXP[r][1]) *
                               dfac(&sd, 2, k, MODE_EXP[r][2]);
                const int64_t node =
                    win_index(lo, n, sd.s[0][i], sd.s[1][j], sd.s[2][k]);
                if (dim < 0) {
                    for (int r = 0; r < nmode; ++r)
                        for (int d = 0; d < ncomp; ++d)
                            cp[r * ncomp + d] += coeff[r] * gu[node * ncomp + d];
                } else {
                    for (int r = 0; r < nmode; ++r)
                        cp[r * ncomp + dim] += coeff[r] * gu[node];
                }
            }
    return ORC_OK;
}

int orc_polypic_g2p(int order, int entity, const double low[3], double h,
                    const int32_t lo[3], const int32_t n[3], int64_t np,
                    const double* x, const double* gu, int ncomp, double* c)
{
    if (ncomp < 1 || ncomp > 3) return ORC_ERR_ARG;
    const int nmode = orc_polypic_nmode(order);
    for (int64_t p = 0; p < np; ++p) {
        int rc = polypic_g2p_one(order, entity, low, h, lo, n, x + 3 * p, gu,
                                 ncomp, c + (int64_t)nmode * ncomp * p);
        if (rc < 0) return rc;
    }
    return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* Velocity gradient, grid normalisation, F update. */

int orc_g2p_gradu(int order, int entity, const double low[3], double h,
                  const int32_t lo[3], const int32_t n[3], int64_t np,
                  const double* x, const double* gu, double* gradu)
{
    if (stag_dim(entity) >= 0) return ORC_ERR_ARG;
    for (int64_t p = 0; p < np; ++p) {
        spl_t sd;
        int nk = spline(order, entity, low, h, x + 3 * p, &sd);
        if (nk < 0) return nk;
        if (!stencil_inside(&sd, lo, n)) return ORC_ERR_BOUNDS;
        double L[9] = {0.0};
        for (int i = 0; i < nk; ++i)
            for (int j = 0; j < nk; ++j)
                for (int k = 0; k < nk; ++k) {
                    const double gw[3] = {sd.g[0][i] * sd.w[1][j] * sd.w[2][k],
                                          sd.w[0][i] * sd.g[1][j] * sd.w[2][k],
                                          sd.w[0][i] * sd.w[1][j] * sd.g[2][k]};
                    const int64_t node =
                        win_index(lo, n, sd.s[0][i], sd.s[1][j], sd.s[2][k]);
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            L[a * 3 + b] += gu[node * 3 + a] * gw[b];
                }
        memcpy(gradu + 9 * p, L, sizeof(L));
    }
    return ORC_OK;
}

/* Generic spline transfers: Picasso::G2P::value / gradient / divergence
 * and Picasso::P2G::value / gradient / divergence
 * (Picasso_ParticleInterpolation.hpp:110-279), which forward to
 * Cabana::Grid::G2P / P2G (Cabana 0.6.1, not in the reference tree).
 * Restated from their definitions with w_n the node weight and gw_n its
 * physical gradient (gw_n[d] = dw/dx_d, products of per-dimension factors
 * as for orc_g2p_gradu):
 *   G2P value       out(p, c)       = sum_n w_n f(n, c)
 *   G2P gradient    out(p, 3c + d)  = sum_n gw_n[d] f(n, c)   (ncomp 1: out(p, d))
 *   G2P divergence  out(p)          = sum_n sum_d gw_n[d] f(n, d)
 *   P2G value       f(n, c)        += w_n v(p, c)
 *   P2G gradient    f(n, d)        += gw_n[d] v(p)
 *   P2G divergence  f(n)           += sum_d gw_n[d] v(p, d)          (vector)
 *                   f(n, i)        += sum_j gw_n[j] v(p, 3i + j)     (tensor)
 * The vector gradient's index order matches orc_g2p_gradu (du_i/dx_j at
 * 3i + j).  Grid fields hold their components contiguously per node. */
static void node_weights(const spl_t* sd, int i, int j, int k, double* w, double gw[3])
{
    *w = sd->w[0][i] * sd->w[1][j] * sd->w[2][k];
    gw[0] = sd->g[0][i] * sd->w[1][j] * sd->w[2][k];
    gw[1] = sd->w[0][i] * sd->g[1][j] * sd->w[2][k];
    gw[2] = sd->w[0][i] * sd->w[1][j] * sd->g[2][k];
}

int orc_g2p_interp(int op, int order, int entity, const double low[3], double h,
                   const int32_t lo[3], const int32_t n[3], int64_t np, const double* x,
                   const double* f, int ncomp, double* out)
{
    int nout;
    if (op == ORC_INTERP_VALUE && ncomp >= 1 && ncomp <= 3) nout = ncomp;
    else if (op == ORC_INTERP_GRADIENT && (ncomp == 1 || ncomp == 3)) nout = 3 * ncomp;
    else if (op == ORC_INTERP_DIVERGENCE && ncomp == 3) nout = 1;
    else return ORC_ERR_ARG;
    for (int64_t p = 0; p < np; ++p) {
        spl_t sd;
        int nk = spline(order, entity, low, h, x + 3 * p, &sd);
        if (nk < 0) return nk;
        if (!stencil_inside(&sd, lo, n)) return ORC_ERR_BOUNDS;
        double r[9] = {0.0};
        for (int i = 0; i < nk; ++i)
            for (int j = 0; j < nk; ++j)
                for (int k = 0; k < nk; ++k) {
                    double w, gw[3];
                    node_weights(&sd, i, j, k, &w, gw);
                    const int64_t node =
                        win_index(lo, n, sd.s[0][i], sd.s[1][j], sd.s[2][k]);
                    const double* fn = f + node * ncomp;
                    for (int c = 0; c < ncomp; ++c) {
                        if (op == ORC_INTERP_VALUE) r[c] += w * fn[c];
                        else if (op == ORC_INTERP_GRADIENT)
                            for (int d = 0; d < 3; ++d) r[3 * c + d] += gw[d] * fn[c];
                        else r[0] += gw[c] * fn[c];
                    }
                }
        memcpy(out + nout * p, r, sizeof(double) * (size_t)nout);
    }
    return ORC_OK;
}

int orc_p2g_interp(int op, int order, int entity, const double low[3], double h,
                   const int32_t lo[3], const int32_t n[3], int64_t np, const double* x,
                   const double* v, int ncomp, double* f)
{
    int ngrid;
    if (op == ORC_INTERP_VALUE && ncomp >= 1 && ncomp <= 3) ngrid = ncomp;
    else if (op == ORC_INTERP_GRADIENT && ncomp == 1) ngrid = 3;
    else if (op == ORC_INTERP_DIVERGENCE && (ncomp == 3 || ncomp == 9)) ngrid = ncomp / 3;
    else return ORC_ERR_ARG;
    for (int64_t p = 0; p < np; ++p) {
        spl_t sd;
        int nk = spline(order, entity, low, h, x + 3 * p, &sd);
        if (nk < 0) return nk;
        if (!stencil_inside(&sd, lo, n)) return ORC_ERR_BOUNDS;
        const double* vp = v + ncomp * p;
        for (int i = 0; i < nk; ++i)
            for (int j = 0; j < nk; ++j)
                for (int k = 0; k < nk; ++k) {
                    double w, gw[3];
                    node_weights(&sd, i, j, k, &w, gw);
                    const int64_t node =
                        win_index(lo, n, sd.s[0][i], sd.s[1][j], sd.s[2][k]);
                    double* fn = f + node * ngrid;
                    if (op == ORC_INTERP_VALUE) {
                        for (int c = 0; c < ncomp; ++c) fn[c] += w * vp[c];
                    } else if (op == ORC_INTERP_GRADIENT) {
                        for (int d = 0; d < 3; ++d) fn[d] += gw[d] * vp[0];
                    } else {
                        for (int a = 0; a < ngrid; ++a) {
                            double s = 0.0;
                            for (int b = 0; b < 3; ++b) s += gw[b] * vp[3 * a + b];
                            fn[a] += s;
                        }
                    }
                }
    }
    return ORC_OK;
}

/* u_i = mu_i / m_i (doc/basic/explict_nearly_incompressible_flow.md:18-23
 * without external forces); nodes with no mass get u = 0. */
int orc_grid_normalize(int64_t nn, int ncomp, const double* m,
                       const double* mu, double* u)
{
    for (int64_t i = 0; i < nn; ++i)
        for (int d = 0; d < ncomp; ++d)
            u[i * ncomp + d] = m[i] > 0.0 ? mu[i * ncomp + d] / m[i] : 0.0;
    return ORC_OK;
}

/* F <- (I + dt*gradu) * F (mat-mat product as in
 * Picasso_BatchedLinearAlgebra.hpp:3806-3834), then det (operator!) and
 * inverse of the new F. */
int orc_update_F(int64_t np, const double* gradu, double dt, double* F,
                 double* det, double* Finv)
{
    for (int64_t p = 0; p < np; ++p) {
        const double* L = gradu + 9 * p;
        double* Fp = F + 9 * p;
        double A[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                A[i * 3 + j] = (i == j ? 1.0 : 0.0) + dt * L[i * 3 + j];
        double Fn[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0.0;
                for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * Fp[k * 3 + j];
                Fn[i * 3 + j] = s;
            }
        memcpy(Fp, Fn, sizeof(Fn));
        det[p] = orc_det3(Fn);
        orc_inverse3(Fn, Finv + 9 * p);
    }
    return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* Cell keys and the stable counting sort. */

int orc_cell_keys(int64_t np, const double* x, const double low[3], double h,
                  const int32_t cell_lo[3], const int32_t ncell[3],
                  const int32_t brick[3], int dual, uint32_t* keys,
                  int32_t* cells)
{
    if (!(h > 0.0)) return ORC_ERR_ARG;
    /* dual: 0 none, 1 every dimension, 0x10 | mask per dimension. */
    int dmask;
    if (dual == 0) dmask = 0;
    else if (dual == 1) dmask = 7;
    else if ((dual & ~7) == 0x10) dmask = dual & 7;
    else return ORC_ERR_ARG;
    int64_t nb[3], bvol = 1, total = 1;
    for (int d = 0; d < 3; ++d) {
        if (ncell[d] <= 0 || brick[d] <= 0) return ORC_ERR_ARG;
        nb[d] = (ncell[d] + brick[d] - 1) / brick[d];
        bvol *= brick[d];
        total *= nb[d];
    }
    if (total * bvol > (int64_t)UINT32_MAX) return ORC_ERR_ARG;
    const double rdx = 1.0 / h;
    for (int64_t p = 0; p < np; ++p) {
        int64_t rel[3];
        for (int d = 0; d < 3; ++d) {
            const double t = (x[3 * p + d] - low[d]) * rdx + (((dmask >> d) & 1) ? 0.5 : 0.0);
            const int64_t c = (int64_t)floor(t);
            if (cells) cells[3 * p + d] = (int32_t)c;
            rel[d] = c - cell_lo[d];
            if (rel[d] < 0 || rel[d] >= ncell[d]) return ORC_ERR_BOUNDS;
        }
        const int64_t bi = rel[0] / brick[0], bj = rel[1] / brick[1],
                      bk = rel[2] / brick[2];
        const int64_t li = rel[0] % brick[0], lj = rel[1] % brick[1],
                      lk = rel[2] % brick[2];
        const int64_t b = (bi * nb[1] + bj) * nb[2] + bk;
        const int64_t l = (li * brick[1] + lj) * brick[2] + lk;
        keys[p] = (uint32_t)(b * bvol + l);
    }
    return ORC_OK;
}

int orc_bin_sort(int64_t np, const uint32_t* keys, uint32_t nkeys,
                 uint32_t* perm, uint32_t* key_start)
{
    memset(key_start, 0, sizeof(uint32_t) * ((size_t)nkeys + 1));
    for (int64_t p = 0; p < np; ++p) {
        if (keys[p] >= nkeys) return ORC_ERR_BOUNDS;
        key_start[keys[p] + 1]++;
    }
    for (uint32_t k = 0; k < nkeys; ++k) key_start[k + 1] += key_start[k];
    uint32_t* cursor = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)nkeys);
    if (!cursor) return ORC_ERR_ALLOC;
    memcpy(cursor, key_start, sizeof(uint32_t) * (size_t)nkeys);
    for (int64_t p = 0; p < np; ++p) perm[cursor[keys[p]]++] = (uint32_t)p;
    free(cursor);
    return ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* CPU baseline step. */

int orc_apic_step_omp(int order, const double low[3], double h,
                      const int32_t lo[3], const int32_t n[3], int64_t np,
                      const double* x, const double* m, double* c,
                      double* gm, double* gmu, double* gu, int mode,
                      int nthreads)
{
    const int64_t nn = (int64_t)n[0] * n[1] * n[2];
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    memset(gm, 0, sizeof(double) * nn);
    memset(gmu, 0, sizeof(double) * nn * 3);
    if (mode == 0) {
        /* Duplicated grids, then a reduction (ScatterDuplicated). */
#pragma omp parallel reduction(| : err)
        {
            double* lm = (double*)calloc((size_t)nn, sizeof(double));
            double* lmu = (double*)calloc((size_t)nn * 3, sizeof(double));
            if (!lm || !lmu) {
                err |= 1;
            } else {
#pragma omp for schedule(static)
                for (int64_t p = 0; p < np; ++p)
                    if (apic_p2g_one(order, ORC_NODE, low, h, lo, n, x + 3 * p,
                                     m[p], c + 12 * p, lm, lmu, 0) < 0)
                        err |= 2;
#pragma omp critical
                {
                    for (int64_t i = 0; i < nn; ++i) gm[i] += lm[i];
                    for (int64_t i = 0; i < 3 * nn; ++i) gmu[i] += lmu[i];
                }
            }
            free(lm);
            free(lmu);
        }
    } else {
#pragma omp parallel for schedule(static) reduction(| : err)
        for (int64_t p = 0; p < np; ++p)
            if (apic_p2g_one(order, ORC_NODE, low, h, lo, n, x + 3 * p, m[p],
                             c + 12 * p, gm, gmu, 1) < 0)
                err |= 2;
    }
    if (err) return (err & 1) ? ORC_ERR_ALLOC : ORC_ERR_BOUNDS;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nn; ++i)
        for (int d = 0; d < 3; ++d)
            gu[i * 3 + d] = gm[i] > 0.0 ? gmu[i * 3 + d] / gm[i] : 0.0;
#pragma omp parallel for schedule(static) reduction(| : err)
    for (int64_t p = 0; p < np; ++p)
        if (apic_g2p_one(order, ORC_NODE, low, h, lo, n, x + 3 * p, gu,
                         c + 12 * p) < 0)
            err |= 2;
    return err ? ORC_ERR_BOUNDS : ORC_OK;
}

/* ---------------------------------------------------------------------- */
/* Particle initialisation (Picasso_ParticleInit.hpp:121-248 random,
 * :279-400 uniform), without the creation functor.  Particle
 * p = cell_id * n + local, cell_id = i + ncell0 (j + ncell1 k) (:170-173);
 * uniform local = ip + d (jp + d kp) (:339-341).  Node i sits at low + i h
 * (LocalMesh::coordinates). */

double orc_uniform01(uint64_t seed, uint64_t pid, int d)
{
    /* SplitMix64 (Steele, Lea, Flood 2014; Vigna's reference code) of the
     * state seed + (3 pid + d + 1) * golden gamma; top 53 bits. */
    uint64_t z = seed + (pid * 3u + (uint64_t)d + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

int64_t orc_init_particles(int mode, const double low[3], double h, const int32_t cell_lo[3],
                           const int32_t ncell[3], int ppc, uint64_t seed, double* x,
                           double* volume)
{
    if (!(h > 0.0) || ppc < 1 || (mode != 0 && mode != 1)) return ORC_ERR_ARG;
    const int64_t npc = mode == 0 ? (int64_t)ppc * ppc * ppc : ppc;
    const int64_t cells = (int64_t)ncell[0] * ncell[1] * ncell[2];
    for (int64_t c = 0; c < cells; ++c) {
        const int32_t own[3] = {(int32_t)(c % ncell[0]), (int32_t)((c / ncell[0]) % ncell[1]),
                                (int32_t)(c / ((int64_t)ncell[0] * ncell[1]))};
        double lc[3], hc[3];
        for (int d = 0; d < 3; ++d) {
            lc[d] = low[d] + (double)(cell_lo[d] + own[d]) * h;
            hc[d] = low[d] + (double)(cell_lo[d] + own[d] + 1) * h;
        }
        for (int64_t l = 0; l < npc; ++l) {
            const int64_t p = c * npc + l;
            if (mode == 0) {
                const int ip[3] = {(int)(l % ppc), (int)((l / ppc) % ppc), (int)(l / (ppc * ppc))};
                for (int d = 0; d < 3; ++d) {
                    const double sp = (hc[d] - lc[d]) / (double)ppc;
                    x[3 * p + d] = 0.5 * sp + (double)ip[d] * sp + lc[d];
                }
            } else {
                for (int d = 0; d < 3; ++d)
                    x[3 * p + d] = lc[d] + (hc[d] - lc[d]) * orc_uniform01(seed, (uint64_t)p, d);
            }
            if (volume) volume[p] = h * h * h / (double)npc;
        }
    }
    return cells * npc;
}

/* ---------------------------------------------------------------------- */
/* Grid halo on one rank, periodic dimensions (the gather / scatter-sum
 * steps of Picasso_GridOperator.hpp:203-215, 465-491).  Window [lo, lo+n),
 * owned [olo, olo+on); node i of a periodic dimension is the image of
 * olo + ((i - olo) mod on).  op 1: scatter-sum, folding one periodic
 * dimension after the other; op 0: gather.  g[(i*n1 + j)*n2 + k][ncomp]. */

static int32_t orc_wrap(int32_t i, int32_t lo, int32_t n)
{
    int32_t r = (i - lo) % n;
    return lo + (r < 0 ? r + n : r);
}

int orc_grid_halo(int op, const int32_t lo[3], const int32_t n[3], int ncomp, double* g,
                  const int32_t olo[3], const int32_t on[3], const int32_t periodic[3])
{
    for (int d = 0; d < 3; ++d)
        if (on[d] < 1 || olo[d] < lo[d] || olo[d] + on[d] > lo[d] + n[d]) return ORC_ERR_ARG;
#define AT(i, j, k) (((int64_t)((i) - lo[0]) * n[1] + ((j) - lo[1])) * n[2] + ((k) - lo[2]))
    if (op == 1) {
        for (int d = 0; d < 3; ++d) {
            if (!periodic[d]) continue;
            for (int32_t i = lo[0]; i < lo[0] + n[0]; ++i)
                for (int32_t j = lo[1]; j < lo[1] + n[1]; ++j)
                    for (int32_t k = lo[2]; k < lo[2] + n[2]; ++k) {
                        int32_t idx[3] = {i, j, k};
                        if (idx[d] >= olo[d] && idx[d] < olo[d] + on[d]) continue;
                        const int64_t src = AT(i, j, k);
                        idx[d] = orc_wrap(idx[d], olo[d], on[d]);
                        const int64_t dst = AT(idx[0], idx[1], idx[2]);
                        for (int c = 0; c < ncomp; ++c) g[dst * ncomp + c] += g[src * ncomp + c];
                    }
        }
    } else {
        for (int32_t i = lo[0]; i < lo[0] + n[0]; ++i)
            for (int32_t j = lo[1]; j < lo[1] + n[1]; ++j)
                for (int32_t k = lo[2]; k < lo[2] + n[2]; ++k) {
                    const int32_t idx[3] = {i, j, k};
                    int32_t img[3];
                    int ghost = 0;
                    for (int d = 0; d < 3; ++d) {
                        img[d] = periodic[d] ? orc_wrap(idx[d], olo[d], on[d]) : idx[d];
                        ghost |= img[d] != idx[d];
                    }
                    if (!ghost) continue;
                    const int64_t src = AT(img[0], img[1], img[2]), dst = AT(i, j, k);
                    for (int c = 0; c < ncomp; ++c) g[dst * ncomp + c] = g[src * ncomp + c];
                }
    }
#undef AT
    return ORC_OK;
}
