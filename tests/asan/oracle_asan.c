/* Memory-safety run of the CPU oracle (oracle/orpic.c): every entry point
 * on small seeded inputs, every buffer allocated at exactly the size the
 * function's contract in orpic.h states, so AddressSanitizer reports any
 * access outside it and UndefinedBehaviorSanitizer any undefined
 * arithmetic.  The transfer cases include particles whose stencils reach
 * the first and the last node of the window.  Built with
 * -fsanitize=address,undefined and run by tests/test_oracle_asan.py; exits
 * 0 when every call returned its expected status. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orpic.h"

static uint64_t rng_state = 0x243F6A8885A308D3ull;

static double urand(void)
{
    /* SplitMix64 step, top 53 bits as a double in [0, 1). */
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* Exactly n doubles (at least one byte), filled with values in [-1, 1). */
static double* drand_alloc(size_t n)
{
    double* p = malloc(n ? n * sizeof(double) : 1);
    if (!p) abort();
    for (size_t i = 0; i < n; ++i) p[i] = 2.0 * urand() - 1.0;
    return p;
}

static double* dzero_alloc(size_t n)
{
    double* p = calloc(n ? n : 1, sizeof(double));
    if (!p) abort();
    return p;
}

static int failures = 0;

#define EXPECT_OK(expr)                                                             \
    do {                                                                            \
        const long long rc_ = (long long)(expr);                                    \
        if (rc_ < 0) {                                                              \
            fprintf(stderr, "%s:%d: %s returned %lld\n", __FILE__, __LINE__, #expr, \
                    rc_);                                                           \
            ++failures;                                                             \
        }                                                                           \
    } while (0)

#define EXPECT_RC(expr, want)                                                          \
    do {                                                                               \
        const long long rc_ = (long long)(expr);                                       \
        if (rc_ != (want)) {                                                           \
            fprintf(stderr, "%s:%d: %s returned %lld, expected %d\n", __FILE__,        \
                    __LINE__, #expr, rc_, (int)(want));                                \
            ++failures;                                                                \
        }                                                                              \
    } while (0)

/* Stencil of xp inside the window [0, n) in every dimension. */
static int stencil_inside(int order, int entity, const double low[3], double h,
                          const int32_t n[3], const double xp[3])
{
    int32_t s[12];
    double w[12], g[12], g2[12], d[12];
    const int nk = orc_spline(order, entity, low, h, xp, s, w, g, g2, d);
    if (nk < 0) return 0;
    for (int dim = 0; dim < 3; ++dim)
        for (int k = 0; k < nk; ++k)
            if (s[4 * dim + k] < 0 || s[4 * dim + k] >= n[dim]) return 0;
    return 1;
}

/* Particles of one (order, entity) case: the ninner of x, then the eight
 * corners of the box of positions whose stencils stay inside the window
 * (per dimension the lowest and the highest such position, in steps of
 * h/64 with the other dimensions at the window centre; the spline is
 * separable, so the corners are inside too).  Returns the count. */
static int64_t case_particles(int order, int entity, const double low[3], double h,
                              const int32_t n[3], const double* x, int64_t ninner, double* xc)
{
    memcpy(xc, x, sizeof(double) * 3 * (size_t)ninner);
    double ext[3][2];
    for (int dim = 0; dim < 3; ++dim)
        for (int end = 0; end < 2; ++end) {
            double xp[3];
            for (int d = 0; d < 3; ++d) xp[d] = low[d] + 0.5 * h * (double)(n[d] - 1);
            const int steps = 64 * (n[dim] - 1);
            for (int t = 0; t <= steps; ++t) {
                const int u = end == 0 ? t : steps - t;
                xp[dim] = low[dim] + h * (double)u / 64.0;
                if (stencil_inside(order, entity, low, h, n, xp)) break;
            }
            ext[dim][end] = xp[dim];
        }
    int64_t np = ninner;
    for (int corner = 0; corner < 8; ++corner, ++np)
        for (int d = 0; d < 3; ++d) xc[3 * np + d] = ext[d][(corner >> d) & 1];
    return np;
}

/* Every transfer of one (order, entity) case on np particles xc, mc. */
static void transfer_case(int order, int entity, const double low[3], double h,
                          const int32_t lo[3], const int32_t n[3], int64_t np, const double* xc,
                          const double* mc)
{
    const size_t nn = (size_t)n[0] * n[1] * n[2];
    const int stag = entity >= ORC_FACE_I;
    const int nc = stag ? 1 : 3;
    int32_t s[12];
    double w[12], g[12], g2[12], dist[12];
    EXPECT_RC(orc_spline(order, entity, low, h, xc, s, w, g, g2, dist), order + 1);

    double* c = drand_alloc(12 * (size_t)np);
    double* gm = dzero_alloc(nn);
    double* gmu = dzero_alloc(nn * nc);
    EXPECT_OK(orc_apic_p2g(order, entity, low, h, lo, n, np, xc, mc, c, gm, gmu));
    EXPECT_OK(orc_apic_g2p(order, entity, low, h, lo, n, np, xc, gmu, c));
    free(c);
    free(gm);
    free(gmu);

    const int nmode = orc_polypic_nmode(order);
    for (int ncomp = 1; ncomp <= 3; ncomp += 2) {
        double* u = drand_alloc((size_t)np * nmode * 3);
        double* cm = drand_alloc((size_t)np * nmode * ncomp);
        double* fmu = dzero_alloc(nn * ncomp);
        double* fm = dzero_alloc(nn);
        EXPECT_OK(orc_polypic_p2g(order, entity, low, h, lo, n, np, xc, mc, u, cm, ncomp, 1e-3,
                                  fmu, fm));
        EXPECT_OK(orc_polypic_g2p(order, entity, low, h, lo, n, np, xc, fmu, ncomp, cm));
        free(u);
        free(cm);
        free(fmu);
        free(fm);
    }

    /* Generic transfers: each (op, ncomp) the contract admits. */
    static const int g2p_cases[][3] = {{ORC_INTERP_VALUE, 1, 1},
                                       {ORC_INTERP_VALUE, 3, 3},
                                       {ORC_INTERP_GRADIENT, 1, 3},
                                       {ORC_INTERP_GRADIENT, 3, 9},
                                       {ORC_INTERP_DIVERGENCE, 3, 1}};
    for (size_t t = 0; t < sizeof(g2p_cases) / sizeof(g2p_cases[0]); ++t) {
        const int op = g2p_cases[t][0], ncomp = g2p_cases[t][1], nout = g2p_cases[t][2];
        double* f = drand_alloc(nn * ncomp);
        double* out = dzero_alloc((size_t)np * nout);
        EXPECT_OK(orc_g2p_interp(op, order, entity, low, h, lo, n, np, xc, f, ncomp, out));
        free(f);
        free(out);
    }
    static const int p2g_cases[][3] = {{ORC_INTERP_VALUE, 2, 2},
                                       {ORC_INTERP_GRADIENT, 1, 3},
                                       {ORC_INTERP_DIVERGENCE, 3, 1},
                                       {ORC_INTERP_DIVERGENCE, 9, 3}};
    for (size_t t = 0; t < sizeof(p2g_cases) / sizeof(p2g_cases[0]); ++t) {
        const int op = p2g_cases[t][0], ncomp = p2g_cases[t][1], ngrid = p2g_cases[t][2];
        double* v = drand_alloc((size_t)np * ncomp);
        double* f = dzero_alloc(nn * ngrid);
        EXPECT_OK(orc_p2g_interp(op, order, entity, low, h, lo, n, np, xc, v, ncomp, f));
        free(v);
        free(f);
    }
    /* An operation the contract rejects touches no buffer. */
    double dummy[1] = {0.0};
    EXPECT_RC(orc_g2p_interp(ORC_INTERP_DIVERGENCE, order, entity, low, h, lo, n, np, xc, dummy,
                             1, dummy),
              ORC_ERR_ARG);

    double* gu = drand_alloc(nn * 3);
    double* gradu = dzero_alloc((size_t)np * 9);
    EXPECT_RC(orc_g2p_gradu(order, entity, low, h, lo, n, np, xc, gu, gradu),
              stag ? ORC_ERR_ARG : ORC_OK);
    free(gu);
    free(gradu);
}

int main(void)
{
    const double h = 0.1;
    const double low[3] = {-0.3, 0.2, 0.05};
    const int32_t lo[3] = {0, 0, 0};
    const int32_t n[3] = {12, 11, 13}; /* node window */
    const size_t nn = (size_t)n[0] * n[1] * n[2];
    const int64_t np = 40;

    /* Particles in the cells [3, n - 4) of every dimension, so that every
     * stencil (orders 1-3, every entity lattice) stays inside the window. */
    double* x = malloc(sizeof(double) * 3 * (size_t)np);
    for (int64_t p = 0; p < np; ++p)
        for (int d = 0; d < 3; ++d)
            x[3 * p + d] = low[d] + h * (3.0 + urand() * (double)(n[d] - 7));
    double* m = drand_alloc((size_t)np + 8);
    for (int64_t p = 0; p < np + 8; ++p) m[p] = 1.0 + fabs(m[p]);

    double* xc = malloc(sizeof(double) * 3 * (size_t)(np + 8));
    for (int order = 1; order <= 3; ++order)
        for (int entity = ORC_NODE; entity <= ORC_EDGE_K; ++entity) {
            const int64_t npc = case_particles(order, entity, low, h, n, x, np, xc);
            transfer_case(order, entity, low, h, lo, n, npc, xc, m);
        }
    free(xc);

    /* A stencil leaving the window is reported, not written. */
    {
        double xo[3] = {low[0] + h * 0.2, low[1] + h * 5.0, low[2] + h * 5.0};
        double mo = 1.0;
        double* c = drand_alloc(12);
        double* gm = dzero_alloc(nn);
        double* gmu = dzero_alloc(nn * 3);
        EXPECT_RC(orc_apic_p2g(2, ORC_NODE, low, h, lo, n, 1, xo, &mo, c, gm, gmu),
                  ORC_ERR_BOUNDS);
        EXPECT_RC(orc_apic_g2p(2, ORC_NODE, low, h, lo, n, 1, xo, gmu, c), ORC_ERR_BOUNDS);
        free(c);
        free(gm);
        free(gmu);
    }

    {
        double* gm = drand_alloc(nn);
        for (size_t i = 0; i < nn; ++i) gm[i] = i % 3 == 0 ? 0.0 : fabs(gm[i]);
        double* mu = drand_alloc(nn * 3);
        double* u = dzero_alloc(nn * 3);
        EXPECT_OK(orc_grid_normalize((int64_t)nn, 3, gm, mu, u));
        free(gm);
        free(mu);
        free(u);
    }

    /* Particle initialisation: exactly the count the mode places. */
    {
        const int32_t cell_lo[3] = {2, 2, 2}, ncell[3] = {3, 4, 2};
        const int64_t ncells = (int64_t)ncell[0] * ncell[1] * ncell[2];
        for (int mode = 0; mode <= 1; ++mode) {
            const int ppc = mode == 0 ? 2 : 5;
            const int64_t want = ncells * (mode == 0 ? ppc * ppc * ppc : ppc);
            double* xi = dzero_alloc((size_t)want * 3);
            double* vol = dzero_alloc((size_t)want);
            EXPECT_RC(orc_init_particles(mode, low, h, cell_lo, ncell, ppc, 7, xi, vol), want);
            free(xi);
            free(vol);
        }
        (void)orc_uniform01(1, 2, 0);
    }

    /* Periodic halo, both operations. */
    {
        const int32_t olo[3] = {2, 2, 2}, on[3] = {n[0] - 4, n[1] - 4, n[2] - 4};
        const int32_t periodic[3] = {1, 0, 1};
        for (int op = 0; op <= 1; ++op) {
            double* gh = drand_alloc(nn * 3);
            EXPECT_OK(orc_grid_halo(op, lo, n, 3, gh, olo, on, periodic));
            free(gh);
        }
    }

    /* Cell keys on three dual modes, then the stable counting sort. */
    {
        const int32_t cell_lo[3] = {1, 1, 1};
        const int32_t ncell[3] = {n[0] - 2, n[1] - 2, n[2] - 2};
        const int32_t brick[3] = {4, 4, 4};
        uint32_t nkeys = 64;
        for (int d = 0; d < 3; ++d) nkeys *= (uint32_t)((ncell[d] + 3) / 4);
        static const int duals[3] = {0, 1, 0x10 | 5};
        for (int t = 0; t < 3; ++t) {
            uint32_t* keys = malloc(sizeof(uint32_t) * (size_t)np);
            int32_t* cells = malloc(sizeof(int32_t) * 3 * (size_t)np);
            uint32_t* perm = malloc(sizeof(uint32_t) * (size_t)np);
            uint32_t* start = malloc(sizeof(uint32_t) * (nkeys + 1));
            EXPECT_OK(orc_cell_keys(np, x, low, h, cell_lo, ncell, brick, duals[t], keys, cells));
            EXPECT_OK(orc_bin_sort(np, keys, nkeys, perm, start));
            free(keys);
            free(cells);
            free(perm);
            free(start);
        }
    }

    /* 3x3 algebra and the deformation-gradient update. */
    {
        double a[9], ai[9];
        for (int i = 0; i < 9; ++i) a[i] = (i % 4 == 0 ? 2.0 : 0.0) + 0.1 * urand();
        (void)orc_det3(a);
        orc_inverse3(a, ai);
        double* gradu = drand_alloc(9 * (size_t)np);
        double* F = dzero_alloc(9 * (size_t)np);
        for (int64_t p = 0; p < np; ++p) F[9 * p] = F[9 * p + 4] = F[9 * p + 8] = 1.0;
        double* det = dzero_alloc((size_t)np);
        double* Finv = dzero_alloc(9 * (size_t)np);
        EXPECT_OK(orc_update_F(np, gradu, 1e-3, F, det, Finv));
        free(gradu);
        free(F);
        free(det);
        free(Finv);
    }

    /* The CPU baseline step, both grid-reduction modes, two threads. */
    for (int mode = 0; mode <= 1; ++mode) {
        double* c = drand_alloc(12 * (size_t)np);
        double* gm = dzero_alloc(nn);
        double* gmu = dzero_alloc(nn * 3);
        double* gu = dzero_alloc(nn * 3);
        EXPECT_OK(orc_apic_step_omp(2, low, h, lo, n, np, x, m, c, gm, gmu, gu, mode, 2));
        free(c);
        free(gm);
        free(gmu);
        free(gu);
    }

    free(x);
    free(m);
    if (failures) {
        fprintf(stderr, "%d unexpected return codes\n", failures);
        return 1;
    }
    printf("oracle asan run: ok\n");
    return 0;
}
