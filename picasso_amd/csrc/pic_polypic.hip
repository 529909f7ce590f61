// PolyPIC particle<->grid transfers (Picasso_PolyPIC.hpp) for gfx950.
//
// The reference defines the linear spline with 8 modes per component
// (basis 1, y0, y1, y2, y0y1, y0y2, y1y2, y0y1y2).  Orders 2 and 3 use 27
// modes, the basis {1, y, y^2}^3: the first eight modes are the linear ones
// in the reference's order, followed by the 19 modes holding a square in
// lexicographic order of their exponents.  g2p mode coefficients are the
// Taylor coefficients of the grid interpolant at the particle: per dimension
// the weight, its gradient, or half its second derivative.
#include "pic_brick.hpp"
#include "pic_common.hpp"
#include "pic_internal.hpp"

namespace pic {

__device__ constexpr int MODE_EXP[27][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1},
    {1, 1, 1}, {0, 0, 2}, {0, 1, 2}, {0, 2, 0}, {0, 2, 1}, {0, 2, 2}, {1, 0, 2},
    {1, 1, 2}, {1, 2, 0}, {1, 2, 1}, {1, 2, 2}, {2, 0, 0}, {2, 0, 1}, {2, 0, 2},
    {2, 1, 0}, {2, 1, 1}, {2, 1, 2}, {2, 2, 0}, {2, 2, 1}, {2, 2, 2}};

template <int ORDER>
struct Modes {
    static constexpr int NM = ORDER == 1 ? 8 : 27;
};

template <typename T>
struct PolyArgs {
    Lattice lat;
    int64_t np;
    const uint32_t* perm;
    const uint32_t* list;  // [count, slots...] to process instead of 0..np-1, or null
    PField<T> x, m, u, c;  // positions, mass, velocity modes (NM x 3), field modes
    GView<T> gmu, gm;
    int pcols;             // columns of the particle field (ncomp)
    int col;               // particle column a staggered transfer works on
    double dt;
    int* err;
};

template <typename T>
__device__ __forceinline__ T ypow(T y, int e)
{
    return e == 0 ? T(1) : (e == 1 ? y : y * y);
}

// PolyPIC::p2g (Picasso_PolyPIC.hpp:39-130, 138-229).  NC is the number of
// grid components (1 for staggered).
template <int ORDER, int NC, bool STAG, typename T>
__device__ __forceinline__ void polypic_p2g_particle(const PolyArgs<T>& a, int64_t p);

// Grid-stride over particles, or over a device-side list of sorted slots
// (list[0] = count) left by the brick kernel.
template <int ORDER, int NC, bool STAG, typename T>
__global__ __launch_bounds__(256) void k_polypic_p2g(PolyArgs<T> a)
{
    const int64_t n = a.list ? int64_t(a.list[0]) : a.np;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t s = a.list ? int64_t(a.list[1 + i]) : i;
        polypic_p2g_particle<ORDER, NC, STAG, T>(a, a.perm ? int64_t(a.perm[s]) : s);
    }
}

template <int ORDER, int NC, bool STAG, typename T>
__device__ __forceinline__ void polypic_p2g_particle(const PolyArgs<T>& a, int64_t p)
{
    constexpr int NK = Knots<ORDER>::NK;
    constexpr int NM = Modes<ORDER>::NM;

    const int64_t xb = a.x.base(p);
    T xp[3] = {a.x.at(xb, 0), a.x.at(xb, 1), a.x.at(xb, 2)};
    const T mp = a.m.at(a.m.base(p), 0);
    Spline<ORDER, T> sd;
    eval_spline<ORDER, T, false, true>(a.lat, xp, sd);
    if (!stencil_inside(sd, a.gmu)) {
        atomicOr(a.err, ERRBIT_BOUNDS);
        return;
    }

    // Affine material motion operator and its inverse (:81-87).
    const int64_t ub = a.u.base(p);
    const T dt = (T)a.dt;
    T am[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int d = 0; d < 3; ++d)
            am[r * 3 + d] = (r == d ? T(1) : T(0)) + dt * a.u.at(ub, (1 + r) * 3 + d);
    T ai[9];
    inverse3(am, det3(am), ai);

    const int64_t cb = a.c.base(p);
    T cm[NM][NC];
#pragma unroll
    for (int r = 0; r < NM; ++r)
#pragma unroll
        for (int d = 0; d < NC; ++d)
            cm[r][d] = STAG ? a.c.at(cb, r * a.pcols + a.col) : a.c.at(cb, r * NC + d);

#pragma unroll
    for (int i = 0; i < NK; ++i)
#pragma unroll
        for (int j = 0; j < NK; ++j)
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                const T wm = mp * sd.w[0][i] * sd.w[1][j] * sd.w[2][k];
                const T dist[3] = {sd.d[0][i], sd.d[1][j], sd.d[2][k]};
                T y[3];
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    y[q] = ai[q * 3 + 0] * dist[0] + ai[q * 3 + 1] * dist[1] + ai[q * 3 + 2] * dist[2];
                T basis[NM];
#pragma unroll
                for (int r = 0; r < NM; ++r)
                    basis[r] = ypow(y[0], MODE_EXP[r][0]) * ypow(y[1], MODE_EXP[r][1]) *
                               ypow(y[2], MODE_EXP[r][2]);
                const int64_t off = a.gmu.off(sd.s0[0] + i, sd.s0[1] + j, sd.s0[2] + k);
#pragma unroll
                for (int d = 0; d < NC; ++d) {
                    // wm * ~c_p.column(d) * basis
                    T v = T(0);
#pragma unroll
                    for (int r = 0; r < NM; ++r) v += (wm * cm[r][d]) * basis[r];
                    atomic_add(a.gmu.data + off + d * a.gmu.cs, v);
                }
                atomic_add(a.gm.data + a.gm.off(sd.s0[0] + i, sd.s0[1] + j, sd.s0[2] + k), wm);
            }
}

// PolyPIC::g2p (Picasso_PolyPIC.hpp:235-293, 299-359).
template <int ORDER, int NC, bool STAG, typename T>
__device__ __forceinline__ void polypic_g2p_particle(const PolyArgs<T>& a, int64_t p);

template <int ORDER, int NC, bool STAG, typename T>
__global__ __launch_bounds__(256) void k_polypic_g2p(PolyArgs<T> a)
{
    const int64_t n = a.list ? int64_t(a.list[0]) : a.np;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t s = a.list ? int64_t(a.list[1 + i]) : i;
        polypic_g2p_particle<ORDER, NC, STAG, T>(a, a.perm ? int64_t(a.perm[s]) : s);
    }
}

template <int ORDER, int NC, bool STAG, typename T>
__device__ __forceinline__ void polypic_g2p_particle(const PolyArgs<T>& a, int64_t p)
{
    constexpr int NK = Knots<ORDER>::NK;
    constexpr int NM = Modes<ORDER>::NM;
    const int64_t xb = a.x.base(p);
    T xp[3] = {a.x.at(xb, 0), a.x.at(xb, 1), a.x.at(xb, 2)};
    Spline<ORDER, T> sd;
    eval_spline<ORDER, T, true, false, (NM > 8)>(a.lat, xp, sd);
    if (!stencil_inside(sd, a.gmu)) {
        atomicOr(a.err, ERRBIT_BOUNDS);
        return;
    }
    T cm[NM][NC];
#pragma unroll
    for (int r = 0; r < NM; ++r)
#pragma unroll
        for (int d = 0; d < NC; ++d) cm[r][d] = T(0);
#pragma unroll
    for (int i = 0; i < NK; ++i)
#pragma unroll
        for (int j = 0; j < NK; ++j)
#pragma unroll
            for (int k = 0; k < NK; ++k) {
                T fac[3][3];
                const int idx[3] = {i, j, k};
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    fac[q][0] = sd.w[q][idx[q]];
                    fac[q][1] = sd.g[q][idx[q]];
                    if constexpr (NM > 8) fac[q][2] = T(0.5) * sd.g2[q][idx[q]];
                    else fac[q][2] = T(0);
                }
                const int64_t off = a.gmu.off(sd.s0[0] + i, sd.s0[1] + j, sd.s0[2] + k);
                T ue[NC];
#pragma unroll
                for (int d = 0; d < NC; ++d) ue[d] = a.gmu.data[off + d * a.gmu.cs];
#pragma unroll
                for (int r = 0; r < NM; ++r) {
                    const T coeff = fac[0][MODE_EXP[r][0]] * fac[1][MODE_EXP[r][1]] *
                                    fac[2][MODE_EXP[r][2]];
#pragma unroll
                    for (int d = 0; d < NC; ++d) cm[r][d] += coeff * ue[d];
                }
            }
    const int64_t cb = a.c.base(p);
    if constexpr (STAG) {
        // Only the working column is reset and rewritten (:328).
#pragma unroll
        for (int r = 0; r < NM; ++r) a.c.at(cb, r * a.pcols + a.col) = cm[r][0];
    } else {
#pragma unroll
        for (int r = 0; r < NM; ++r)
#pragma unroll
            for (int d = 0; d < NC; ++d) a.c.at(cb, r * NC + d) = cm[r][d];
    }
}

// ---------------------------------------------------------------------------
// Brick p2g (collocated, three grid components) for particles binned by
// pic_bin_sort on the spline's knot-base lattice: the APIC brick scheme
// with the PolyPIC node value.  Four work-items per bin of a 4^3 brick, one
// per grid component (mass, momentum x/y/z), each walking every particle of
// its bin and accumulating its component at the NK^3 stencil nodes in
// registers; the bin sums go through LDS to one atomic per footprint node
// (gather_flush).  Particles are staged per chunk with coalesced loads:
// (f0, f1, f2, m) and the rows of A^-1 (A = I + dt U, Picasso_PolyPIC.hpp:
// 81-87), computed once per particle.  A work-item reads its component's
// modes from memory and evaluates sum_r c(r, d) b_r(y) at y = A^-1 d_n as a
// nested Horner polynomial; y is separable in the node offsets,
// y = P_i + Q_j + R_k.
template <typename T>
struct PolyBrickArgs {
    PolyArgs<T> a;
    PIC_BRICK_WALK_FIELDS
};

template <int ORDER, typename T>
struct PolyBrickShape {
    static constexpr int PB = 4;
    static constexpr int NK = ORDER + 1;
    static constexpr int NN = NK * NK * NK;
    static constexpr int BINS = PB * PB * PB;
    static constexpr int NT = BINS * 4;
    static constexpr int SLOTS = NN * BINS;
    static constexpr int SUM_BUDGET = sizeof(T) == 4 ? 34 * 1024 : 17 * 1024;
    static constexpr int NROUND = 4 * SLOTS * int(sizeof(T)) <= SUM_BUDGET   ? 1
                                  : 2 * SLOTS * int(sizeof(T)) <= SUM_BUDGET ? 2
                                                                              : 4;
    static constexpr int CPR = 4 / NROUND;
    static constexpr int CAP = sizeof(T) == 4 ? 576 : 200;
    static constexpr int STAGE = CAP * 16;
    static constexpr int SUMS = CPR * SLOTS;
    static constexpr int LDS = STAGE > SUMS ? STAGE : SUMS;
    // Waves per SIMD the registers are held to: the cubic FP32 kernel fits
    // two in 256 VGPRs without spilling; FP64 would spill, so it runs one.
    static constexpr int WAVES = sizeof(T) == 4 ? 2 : 1;
};

// sum over exponents e of C[e0][e1][e2] y0^e0 y1^e1 y2^e2 (E exponents per
// dimension), innermost dimension first.
template <int E, typename T>
__device__ __forceinline__ T nested_horner(const T (&C)[E][E][E], const T y[3])
{
    T mid[E];
#pragma unroll
    for (int e0 = 0; e0 < E; ++e0) {
        T inner[E];
#pragma unroll
        for (int e1 = 0; e1 < E; ++e1) {
            T v = C[e0][e1][E - 1];
#pragma unroll
            for (int e2 = E - 2; e2 >= 0; --e2) v = v * y[2] + C[e0][e1][e2];
            inner[e1] = v;
        }
        T v = inner[E - 1];
#pragma unroll
        for (int e1 = E - 2; e1 >= 0; --e1) v = v * y[1] + inner[e1];
        mid[e0] = v;
    }
    T v = mid[E - 1];
#pragma unroll
    for (int e0 = E - 2; e0 >= 0; --e0) v = v * y[0] + mid[e0];
    return v;
}

template <int ORDER, typename T>
using PolyStage = BrickStage<(PolyBrickShape<ORDER, T>::CAP + PolyBrickShape<ORDER, T>::NT - 1) /
                                 PolyBrickShape<ORDER, T>::NT,
                             PolyBrickShape<ORDER, T>::NT, PolyBrickShape<ORDER, T>::BINS>;

// Stage chunk [c0, c1) of a brick's sorted slots for the PolyPIC p2g
// bodies: per particle (f, m) and the rows of A^-1 (A = I + dt U, inverted
// once here) in the stage_at groups, its index in pidx, its bin's
// chunk-local range in the returned range buffer; a particle whose knot
// base differs from its bin is pushed to the stale list (f0 = STALE_F).
template <int ORDER, typename T>
__device__ __forceinline__ uint32_t* polypic_stage_chunk(const PolyBrickArgs<T>& g,
                                                         PolyStage<ORDER, T>& st, uint32_t beg,
                                                         uint32_t c0, uint32_t c1, uint32_t kb,
                                                         const int f0[3], T* lds,
                                                         uint32_t* pidx)
{
    using S = PolyBrickShape<ORDER, T>;
    using T4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    constexpr int PB = S::PB, BINS = S::BINS, NT = S::NT, CAP = S::CAP;
    constexpr int FIRST = Knots<ORDER>::FIRST;
    constexpr int K = (CAP + NT - 1) / NT;
    const PolyArgs<T>& a = g.a;
    const T dt = (T)a.dt;
    uint32_t* rng = st.begin_chunk(g, a.perm, beg, c0, c1);
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const uint32_t s = c0 + threadIdx.x + uint32_t(k * NT);
        if (s >= c1) continue;
        const int64_t p = int64_t(st.cs.perm[k]);
        const int q = int(s - c0);
        const int lb = int(st.cs.key[k] - kb);
        const int eb[3] = {f0[0] + lb / (PB * PB), f0[1] + (lb / PB) % PB, f0[2] + lb % PB};
        const int64_t xb = a.x.base(p);
        T4 v;
        T* fr = &v.x;
        bool ok = true;
#pragma unroll
        for (int dim = 0; dim < 3; ++dim) {
            double fd;
            const double kb0 = knot_base<ORDER>(double(a.x.at(xb, dim)), a.lat.o[dim],
                                                 a.lat.rdx, fd);
            ok = ok && int(kb0) + FIRST == eb[dim];
            fr[dim] = T(fd);
        }
        if (!ok) {
            push_stale(g, s);
            v.x = T(STALE_F);
        }
        v.w = a.m.at(a.m.base(p), 0);
        *reinterpret_cast<T4*>(lds + stage_at(q, 0)) = v;
        // A = I + dt U, U(r, d) = u_p(1 + r, d) (:81-84); rows of A^-1.
        const int64_t ub = a.u.base(p);
        T am[9], ai[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int e = 0; e < 3; ++e)
                am[r * 3 + e] = (r == e ? T(1) : T(0)) + dt * a.u.at(ub, (1 + r) * 3 + e);
        inverse3(am, det3(am), ai);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            v.x = ai[r * 3 + 0];
            v.y = ai[r * 3 + 1];
            v.z = ai[r * 3 + 2];
            v.w = T(0);
            *reinterpret_cast<T4*>(lds + stage_at(q, 1 + r)) = v;
        }
        pidx[q] = uint32_t(p);
        st.cs.template note<BINS>(k, q, kb, rng);
    }
    st.end_chunk();
    return rng;
}

template <int ORDER, typename T, bool STAG>
__device__ __forceinline__ void polypic_p2g_one_brick(const PolyBrickArgs<T>& g,
                                                      const BrickItem& it,
                                                      BrickWalk<true, PolyBrickArgs<T>>& w,
                                                      T* lds, uint32_t* pidx,
                                                      PolyStage<ORDER, T>& st, float* mbuf)
{
    using S = PolyBrickShape<ORDER, T>;
    using T4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    constexpr int PB = S::PB, NK = S::NK, NN = S::NN, BINS = S::BINS, NT = S::NT, CAP = S::CAP;
    constexpr int NM = Modes<ORDER>::NM;
    constexpr int E = ORDER == 1 ? 2 : 3;
    const PolyArgs<T>& a = g.a;
    const uint32_t beg = it.beg, end = it.end;
    int f0[3];
    brick_origin<PB>(g, it.b, f0);
    const uint32_t kb = it.b * BINS;

    const int l = int(threadIdx.x) >> 2, comp = int(threadIdx.x) & 3;
    const int d = comp > 0 ? comp - 1 : 0;
    // Cubic, three columns (SLAB): the four work-items of a bin split the
    // 4^3 stencil by its slab i = comp instead of by grid component; each
    // sums mass and the three momentum components at its 16 nodes.  The
    // offsets y = A^-1 d_n of a node are then formed once for the three
    // components, and no work-item spends the 64-node mode sums on the
    // mass (whose coefficients are the constant 1): 29 % fewer vector
    // instructions per particle than one component per work-item.
    constexpr bool SLAB = ORDER == 3 && !STAG;
    constexpr int NA = SLAB ? NK * NK : NN;  // nodes per work-item
    constexpr int CA = SLAB ? 4 : 1;         // components per node
    T acc[NA][CA];
#pragma unroll
    for (int n = 0; n < NA; ++n)
#pragma unroll
        for (int cc = 0; cc < CA; ++cc) acc[n][cc] = T(0);
    for (uint32_t c0 = beg; c0 < end; c0 += CAP) {
        const uint32_t c1 = min(end, c0 + uint32_t(CAP));
        uint32_t* rng = polypic_stage_chunk<ORDER, T>(g, st, beg, c0, c1, kb, f0, lds, pidx);
        __syncthreads();
        if (c0 == beg) w.fetch();
        const uint32_t q0 = rng[l], q1 = rng[BINS + l];
        if constexpr (SLAB) {
            const int il = comp;
            // FP32 (MDMA): the 81 modes of a bin's next particle are copied
            // from HBM into the wave's LDS buffer while the current one is
            // summed, without registers: copy k takes mode 4 k + comp of the
            // bin's particle (lane 4 l + comp of the wave), so the bin's
            // work-items read modes 4 k .. 4 k + 3 back with one 16-byte
            // load.  Otherwise each work-item loads the modes from HBM.
            constexpr bool MDMA = sizeof(T) == 4;
            constexpr int NCP = (3 * NM + 3) / 4;  // copies per particle
            float* wbuf = mbuf + __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6) * (NCP * 64);
            const int lw = (int(threadIdx.x) & 63) >> 2;  // the bin's position in the wave
            auto copy_modes = [&](uint32_t qn) {
                const T* src = &a.c.at(a.c.base(int64_t(pidx[qn])), comp);
                const int64_t step = 4 * a.c.cs;
#pragma unroll
                for (int k = 0; k < NCP; ++k) {
                    if (4 * k + comp < 3 * NM) load_to_lds(src, wbuf + 64 * k);
                    src += step;
                }
            };
            if constexpr (MDMA)
                if (q0 < q1) copy_modes(q0);
            for (uint32_t qq = q0; qq < q1; ++qq) {
                const int q = int(qq);
                T C[3][E][E][E];
                if constexpr (MDMA) {
                    __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);  // this particle's modes
#pragma unroll
                    for (int k = 0; k < NCP; ++k) {
                        const float4 v = *reinterpret_cast<const float4*>(wbuf + 64 * k + 4 * lw);
                        const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                        for (int cc = 0; cc < 4; ++cc) {
                            const int e = 4 * k + cc;
                            if (e < 3 * NM)
                                C[e % 3][MODE_EXP[e / 3][0]][MODE_EXP[e / 3][1]][MODE_EXP[e / 3][2]] =
                                    T(vv[cc]);
                        }
                    }
                    // the buffer is read before the next particle's copies
                    // land in it
                    __builtin_amdgcn_s_waitcnt(WAIT_LGKMCNT0);
                    if (qq + 1 < q1) copy_modes(qq + 1);
                }
                const T4 fm = *reinterpret_cast<const T4*>(lds + stage_at(q, 0));
                if (fm.x == T(STALE_F)) continue;
                T ai[3][3];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const T4 row = *reinterpret_cast<const T4*>(lds + stage_at(q, 1 + r));
                    ai[r][0] = row.x;
                    ai[r][1] = row.y;
                    ai[r][2] = row.z;
                }
                // the modes of all three columns of c_p
                if constexpr (!MDMA) {
                    const int64_t cb = a.c.base(int64_t(pidx[q]));
#pragma unroll
                    for (int r = 0; r < NM; ++r)
#pragma unroll
                        for (int dd = 0; dd < 3; ++dd)
                            C[dd][MODE_EXP[r][0]][MODE_EXP[r][1]][MODE_EXP[r][2]] =
                                a.c.at(cb, r * 3 + dd);
                }
                const T fr[3] = {fm.x, fm.y, fm.z};
                Spline<ORDER, T> sd;
                spline_from_frac<ORDER, T, false, true>(a.lat, fr, sd);
                // this slab's distance and weight along dimension 0
                // (selected, not indexed: il varies across lanes)
                T d0 = sd.d[0][0], w0 = sd.w[0][0];
#pragma unroll
                for (int n = 1; n < NK; ++n) {
                    d0 = il == n ? sd.d[0][n] : d0;
                    w0 = il == n ? sd.w[0][n] : w0;
                }
                // y = A^-1 d_n = P_i + Q_j + R_k.
                T P[3], Q[NK][3], R[NK][3];
#pragma unroll
                for (int r = 0; r < 3; ++r) P[r] = ai[r][0] * d0;
#pragma unroll
                for (int n = 0; n < NK; ++n)
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        Q[n][r] = ai[r][1] * sd.d[1][n];
                        R[n][r] = ai[r][2] * sd.d[2][n];
                    }
                const T wim = fm.w * w0;
#pragma unroll
                for (int j = 0; j < NK; ++j) {
                    const T wij = wim * sd.w[1][j];
                    const T pq[3] = {P[0] + Q[j][0], P[1] + Q[j][1], P[2] + Q[j][2]};
#pragma unroll
                    for (int k = 0; k < NK; ++k) {
                        const T y[3] = {pq[0] + R[k][0], pq[1] + R[k][1], pq[2] + R[k][2]};
                        const T wn = wij * sd.w[2][k];
                        T* an = acc[j * NK + k];
                        an[0] += wn;
#pragma unroll
                        for (int dd = 0; dd < 3; ++dd) an[1 + dd] += wn * nested_horner<E>(C[dd], y);
                    }
                }
            }
            __syncthreads();
            continue;
        }
        for (uint32_t qq = q0; qq < q1; ++qq) {
            const int q = int(qq);
            const T4 fm = *reinterpret_cast<const T4*>(lds + stage_at(q, 0));
            if (fm.x == T(STALE_F)) continue;
            T ai[3][3];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const T4 row = *reinterpret_cast<const T4*>(lds + stage_at(q, 1 + r));
                ai[r][0] = row.x;
                ai[r][1] = row.y;
                ai[r][2] = row.z;
            }
            // Coefficients of this work-item's component: the modes of
            // column d of c_p, or the constant 1 for the mass.
            T C[E][E][E];
#pragma unroll
            for (int e0 = 0; e0 < E; ++e0)
#pragma unroll
                for (int e1 = 0; e1 < E; ++e1)
#pragma unroll
                    for (int e2 = 0; e2 < E; ++e2) C[e0][e1][e2] = T(0);
            const int64_t cb = a.c.base(int64_t(pidx[q]));
#pragma unroll
            for (int r = 0; r < NM; ++r) {
                // Staggered: every momentum work-item takes the one column
                // the transfer works on; only its sum is written.
                const T cr = STAG ? a.c.at(cb, r * a.pcols + a.col) : a.c.at(cb, r * 3 + d);
                C[MODE_EXP[r][0]][MODE_EXP[r][1]][MODE_EXP[r][2]] =
                    comp == 0 ? (r == 0 ? T(1) : T(0)) : cr;
            }
            const T fr[3] = {fm.x, fm.y, fm.z};
            Spline<ORDER, T> sd;
            spline_from_frac<ORDER, T, false, true>(a.lat, fr, sd);
            // y = A^-1 d_n = P_i + Q_j + R_k.
            T P[NK][3], Q[NK][3], R[NK][3];
#pragma unroll
            for (int n = 0; n < NK; ++n)
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    P[n][r] = ai[r][0] * sd.d[0][n];
                    Q[n][r] = ai[r][1] * sd.d[1][n];
                    R[n][r] = ai[r][2] * sd.d[2][n];
                }
#pragma unroll
            for (int i = 0; i < NK; ++i) {
                const T wim = fm.w * sd.w[0][i];
#pragma unroll
                for (int j = 0; j < NK; ++j) {
                    const T wij = wim * sd.w[1][j];
                    const T pq[3] = {P[i][0] + Q[j][0], P[i][1] + Q[j][1], P[i][2] + Q[j][2]};
#pragma unroll
                    for (int k = 0; k < NK; ++k) {
                        const T y[3] = {pq[0] + R[k][0], pq[1] + R[k][1], pq[2] + R[k][2]};
                        acc[(i * NK + j) * NK + k][0] += (wij * sd.w[2][k]) * nested_horner<E>(C, y);
                    }
                }
            }
        }
        __syncthreads();
    }

    for (int r = 0; r < S::NROUND; ++r) {
        const int r_lo = r * S::CPR;
        if constexpr (SLAB) {
            // this work-item's slab: offsets (comp, j, k), components
            // r_lo .. r_lo + CPR
            T* dst = lds + l * S::CPR;
#pragma unroll
            for (int n = 0; n < NA; ++n)
#pragma unroll
                for (int cc = 0; cc < S::CPR; ++cc)
                    dst[(comp * NA + n) * BINS * S::CPR + cc] = acc[n][r_lo + cc];
        } else if (comp >= r_lo && comp < r_lo + S::CPR) {
            T* dst = lds + l * S::CPR + (comp - r_lo);
#pragma unroll
            for (int n = 0; n < NN; ++n) dst[n * BINS * S::CPR] = acc[n][0];
        }
        if (r == 0) w.publish();
        __syncthreads();
        if (r == 0) st.prefetch_next(g, a.perm, w, CAP);
        gather_flush<NK, PB, NT, S::CPR>(lds, f0, r_lo, 1, a.gm, a.gmu, a.err, STAG ? 2 : 4);
        if (r + 1 < S::NROUND) __syncthreads();
    }
}

// Brick g2p (three components, or the one of a staggered field): the
// brick's (4+ORDER)^3 grid footprint is staged in LDS once, then one
// work-item per particle forms its modes, the Taylor coefficients
// sum_n F0_e0(i) F1_e1(j) F2_e2(k) u_n (F_e = weight, gradient, half second
// derivative), summed one dimension at a time (k, then j, then i) instead
// of per node and mode.  A staggered g2p rewrites only its column
// (Picasso_PolyPIC.hpp:328).
template <int ORDER, typename T, bool STAG>
__device__ __forceinline__ void polypic_g2p_one_brick(const PolyBrickArgs<T>& g,
                                                      const BrickItem& it,
                                                      BrickWalk<true, PolyBrickArgs<T>>& w,
                                                      T* us)
{
    constexpr int PB = 4, NT = 256;
    constexpr int NK = ORDER + 1;
    constexpr int FN = PB + ORDER;
    constexpr int FV = FN * FN * FN;
    constexpr int NM = Modes<ORDER>::NM;
    constexpr int E = ORDER == 1 ? 2 : 3;
    constexpr int NC = STAG ? 1 : 3;
    const PolyArgs<T>& a = g.a;
    const uint32_t beg = it.beg, end = it.end;
    int f0[3];
    brick_origin<PB>(g, it.b, f0);
    const GView<T>& gu = a.gmu;
    w.fetch();
    for (int q = threadIdx.x; q < FV; q += NT) {
        const int ni = f0[0] + q / (FN * FN), nj = f0[1] + (q / FN) % FN, nk = f0[2] + q % FN;
        const bool in = gu.covers(0, ni, ni) && gu.covers(1, nj, nj) && gu.covers(2, nk, nk);
        const int64_t off = in ? gu.off(ni, nj, nk) : 0;
#pragma unroll
        for (int d = 0; d < NC; ++d) us[d * FV + q] = in ? gu.data[off + d * gu.cs] : T(0);
    }
    w.publish();
    __syncthreads();
    for (uint32_t s = beg + threadIdx.x; s < end; s += NT) {
        const int64_t p = int64_t(a.perm[s]);
        const int64_t xb = a.x.base(p);
        const T xp[3] = {a.x.at(xb, 0), a.x.at(xb, 1), a.x.at(xb, 2)};
        Spline<ORDER, T> sd;
        eval_spline<ORDER, T, true, false, (NM > 8)>(a.lat, xp, sd);
        const int li = sd.s0[0] - f0[0], lj = sd.s0[1] - f0[1], lk = sd.s0[2] - f0[2];
        if (!(li >= 0 && li < PB && lj >= 0 && lj < PB && lk >= 0 && lk < PB)) {
            push_stale(g, s);
            continue;
        }
        // Footprint nodes outside the window were staged as zeros: a
        // stencil reaching them is an error, as in the per-particle gather.
        if (!stencil_inside(sd, gu)) {
            atomicOr(a.err, ERRBIT_BOUNDS);
            continue;
        }
        // F[q][e](n): weight, gradient, half second derivative of dim q.
        T F[3][E][NK];
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int n = 0; n < NK; ++n) {
                F[q][0][n] = sd.w[q][n];
                F[q][1][n] = sd.g[q][n];
                if constexpr (E > 2) F[q][2][n] = T(0.5) * sd.g2[q][n];
            }
        const int lbase = (li * FN + lj) * FN + lk;
        const int64_t cb = a.c.base(p);
        // One grid component at a time: its 27 sums in registers (the three
        // at once held 190 VGPRs, two waves per SIMD; the loads and the
        // arithmetic are the same either way).
#pragma unroll 1
        for (int d = 0; d < NC; ++d) {
            const T* ud = us + d * FV + lbase;
            T cm[E][E][E];
#pragma unroll
            for (int e0 = 0; e0 < E; ++e0)
#pragma unroll
                for (int e1 = 0; e1 < E; ++e1)
#pragma unroll
                    for (int e2 = 0; e2 < E; ++e2) cm[e0][e1][e2] = T(0);
#pragma unroll
            for (int i = 0; i < NK; ++i) {
                T t[E][E];
#pragma unroll
                for (int e1 = 0; e1 < E; ++e1)
#pragma unroll
                    for (int e2 = 0; e2 < E; ++e2) t[e1][e2] = T(0);
#pragma unroll
                for (int j = 0; j < NK; ++j) {
                    T sk[E];
#pragma unroll
                    for (int e2 = 0; e2 < E; ++e2) sk[e2] = T(0);
#pragma unroll
                    for (int k = 0; k < NK; ++k) {
                        const T ue = ud[(i * FN + j) * FN + k];
#pragma unroll
                        for (int e2 = 0; e2 < E; ++e2) sk[e2] += F[2][e2][k] * ue;
                    }
#pragma unroll
                    for (int e1 = 0; e1 < E; ++e1)
#pragma unroll
                        for (int e2 = 0; e2 < E; ++e2) t[e1][e2] += F[1][e1][j] * sk[e2];
                }
#pragma unroll
                for (int e0 = 0; e0 < E; ++e0)
#pragma unroll
                    for (int e1 = 0; e1 < E; ++e1)
#pragma unroll
                        for (int e2 = 0; e2 < E; ++e2) cm[e0][e1][e2] += F[0][e0][i] * t[e1][e2];
            }
#pragma unroll
            for (int r = 0; r < NM; ++r) {
                const T ce = cm[MODE_EXP[r][0]][MODE_EXP[r][1]][MODE_EXP[r][2]];
                if constexpr (STAG) a.c.at(cb, r * a.pcols + a.col) = ce;
                else a.c.at(cb, r * 3 + d) = ce;
            }
        }
    }
}

template <int ORDER, typename T, bool STAG>
__global__ __launch_bounds__(256) void k_polypic_g2p_brick(PolyBrickArgs<T> g)
{
    constexpr int FN = 4 + ORDER;
    __shared__ T us[(STAG ? 1 : 3) * FN * FN * FN];
    __shared__ BrickItem q[2];
    walk_bricks<true>(g, q, [&](const BrickItem& it, BrickWalk<true, PolyBrickArgs<T>>& w) {
        polypic_g2p_one_brick<ORDER, T, STAG>(g, it, w, us);
    });
}

template <int ORDER, typename T, bool STAG>
__global__ __launch_bounds__((PolyBrickShape<ORDER, T>::NT), (PolyBrickShape<ORDER, T>::WAVES)) void k_polypic_p2g_brick(
    PolyBrickArgs<T> g)
{
    using S = PolyBrickShape<ORDER, T>;
    __shared__ __align__(16) T lds[S::LDS];
    __shared__ uint32_t pidx[S::CAP];
    __shared__ uint32_t rng[2 * 2 * S::BINS];
    __shared__ BrickItem q[2];
    // FP32 cubic three-column p2g: per wave, the LDS copies of one
    // particle's 81 modes for each of its 16 bins
    constexpr int MB = ORDER == 3 && !STAG && sizeof(T) == 4 ? (S::NT / 64) * 21 * 64 : 4;
    __shared__ __align__(16) float mbuf[MB];
    for (int i = threadIdx.x; i < 4 * S::BINS; i += S::NT) rng[i] = 0u;
    PolyStage<ORDER, T> st{rng, 0, false, {}};
    walk_bricks<true>(g, q, [&](const BrickItem& it, BrickWalk<true, PolyBrickArgs<T>>& w) {
        polypic_p2g_one_brick<ORDER, T, STAG>(g, it, w, lds, pidx, st, mbuf);
    });
}

template <typename T>
static int polypic_p2g_brick_run(pic_context* ctx, int order, int entity, int ncomp,
                                 const pic_mesh& mesh, int64_t np, const pic_pfield* x,
                                 const pic_pfield* m, const pic_pfield* u, const pic_pfield* c,
                                 const pic_grid& gmu, const pic_grid* gm, double dt,
                                 const pic_bins& bins);

template <typename T>
struct PolyLaunch {
    PolyArgs<T> a;
    hipStream_t st;
    bool p2g;
    template <int O, int NC, bool S>
    void go()
    {
        const unsigned b = a.list ? 512u : blocks_for(a.np, 256);
        if (p2g) k_polypic_p2g<O, NC, S, T><<<b, 256, 0, st>>>(a);
        else k_polypic_g2p<O, NC, S, T><<<b, 256, 0, st>>>(a);
    }
    template <int O>
    void by_nc(int nc, bool stag)
    {
        if (stag) go<O, 1, true>();
        else if (nc == 1) go<O, 1, false>();
        else if (nc == 2) go<O, 2, false>();
        else go<O, 3, false>();
    }
    void run(int order, int nc, bool stag)
    {
        if (order == 1) by_nc<1>(nc, stag);
        else if (order == 2) by_nc<2>(nc, stag);
        else by_nc<3>(nc, stag);
    }
};

template <typename T>
static PolyArgs<T> make_poly_args(pic_context* ctx, int entity, int ncomp, const pic_mesh& mesh,
                                  int64_t np, const pic_pfield* x, const pic_pfield* m,
                                  const pic_pfield* u, const pic_pfield* c, const pic_grid& gmu,
                                  const pic_grid* gm, double dt, const uint32_t* perm)
{
    PolyArgs<T> a;
    a.lat = make_lattice(mesh, entity);
    a.np = np;
    a.perm = perm;
    a.list = nullptr;
    a.x = make_pfield<T>(*x);
    a.m = m ? make_pfield<T>(*m) : PField<T>{};
    a.u = u ? make_pfield<T>(*u) : PField<T>{};
    a.c = make_pfield<T>(*c);
    a.gmu = make_gview<T>(gmu);
    a.gm = gm ? make_gview<T>(*gm) : GView<T>{};
    const int sd = stag_dim(entity);
    // Staggered: the column is entity dim for a 3-column particle field,
    // else 0 (:181, :325).
    a.pcols = ncomp;
    a.col = (sd >= 0 && ncomp == 3) ? sd : 0;
    a.dt = dt;
    a.err = ctx->d_err;
    return a;
}

// Brick paths for a three-column collocated field or any staggered one
// (ncomp is the particle field's column count).
template <typename T>
static int polypic_p2g_brick_run(pic_context* ctx, int order, int entity, int ncomp,
                                 const pic_mesh& mesh, int64_t np, const pic_pfield* x,
                                 const pic_pfield* m, const pic_pfield* u, const pic_pfield* c,
                                 const pic_grid& gmu, const pic_grid* gm, double dt,
                                 const pic_bins& bins)
{
    if (brick_edge(bins.geom) != 4) return PIC_ERR_UNSUPPORTED;
    PolyBrickArgs<T> g;
    if (const int rc = setup_walk(ctx, order, entity, np, bins, g)) return rc;
    g.a = make_poly_args<T>(ctx, entity, ncomp, mesh, np, x, m, u, c, gmu, gm, dt, bins.perm);
    const bool stag = stag_dim(entity) >= 0;
    {
        TimedLaunch tl(ctx, "polypic_p2g");
        auto go = [&](auto kernel, int nt) {
            const unsigned nbk = persistent_blocks(ctx, kernel, nt, 0, g.nbricks);
            kernel<<<nbk, nt, 0, ctx->stream>>>(g);
        };
        if (stag) {
            if (order == 1) go(k_polypic_p2g_brick<1, T, true>, PolyBrickShape<1, T>::NT);
            else if (order == 2) go(k_polypic_p2g_brick<2, T, true>, PolyBrickShape<2, T>::NT);
            else go(k_polypic_p2g_brick<3, T, true>, PolyBrickShape<3, T>::NT);
        } else {
            if (order == 1) go(k_polypic_p2g_brick<1, T, false>, PolyBrickShape<1, T>::NT);
            else if (order == 2) go(k_polypic_p2g_brick<2, T, false>, PolyBrickShape<2, T>::NT);
            else go(k_polypic_p2g_brick<3, T, false>, PolyBrickShape<3, T>::NT);
        }
        if (launch_status()) return PIC_ERR_HIP;
    }
    // Particles the brick kernel skipped (stale bins).
    TimedLaunch tl(ctx, "polypic_p2g_stale");
    PolyLaunch<T> L{g.a, ctx->stream, true};
    L.a.list = g.stale;
    L.run(order, ncomp, stag);
    return launch_status();
}

template <typename T>
static int polypic_g2p_brick_run(pic_context* ctx, int order, int entity, int ncomp,
                                 const pic_mesh& mesh, int64_t np, const pic_pfield* x,
                                 const pic_pfield* c, const pic_grid& gu, const pic_bins& bins)
{
    if (brick_edge(bins.geom) != 4) return PIC_ERR_UNSUPPORTED;
    PolyBrickArgs<T> g;
    if (const int rc = setup_walk(ctx, order, entity, np, bins, g)) return rc;
    g.a = make_poly_args<T>(ctx, entity, ncomp, mesh, np, x, nullptr, nullptr, c, gu, nullptr,
                            0.0, bins.perm);
    const bool stag = stag_dim(entity) >= 0;
    {
        TimedLaunch tl(ctx, "polypic_g2p");
        auto go = [&](auto kernel) {
            const unsigned nbk = persistent_blocks(ctx, kernel, 256, 0, g.nbricks);
            kernel<<<nbk, 256, 0, ctx->stream>>>(g);
        };
        if (stag) {
            if (order == 1) go(k_polypic_g2p_brick<1, T, true>);
            else if (order == 2) go(k_polypic_g2p_brick<2, T, true>);
            else go(k_polypic_g2p_brick<3, T, true>);
        } else {
            if (order == 1) go(k_polypic_g2p_brick<1, T, false>);
            else if (order == 2) go(k_polypic_g2p_brick<2, T, false>);
            else go(k_polypic_g2p_brick<3, T, false>);
        }
        if (launch_status()) return PIC_ERR_HIP;
    }
    TimedLaunch tl(ctx, "polypic_g2p_stale");
    PolyLaunch<T> L{g.a, ctx->stream, false};
    L.a.list = g.stale;
    L.run(order, ncomp, stag);
    return launch_status();
}

}  // namespace pic

using namespace pic;

extern "C" int pic_polypic_nmode(int order)
{
    if (order < 1 || order > 3) return PIC_ERR_ARG;
    return order == 1 ? 8 : 27;
}

extern "C" int pic_polypic_p2g(pic_context* ctx, int order, int entity, int dtype,
                               const pic_mesh* mesh, int64_t np, int ncomp,
                               const pic_pfield* x, const pic_pfield* m_p,
                               const pic_pfield* u_p, const pic_pfield* c_p,
                               const pic_grid* mu_i, const pic_grid* m_i, double dt,
                               const pic_bins* bins, uint32_t flags)
{
    Region region("picasso_amd::PolyPIC::p2g");
    if (!ctx || !mesh || !(mesh->h > 0.0) || order < 1 || order > 3 ||
        !valid_entity(entity) || (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 ||
        np > int64_t(UINT32_MAX) || ncomp < 1 || ncomp > 3 || !valid_grid(mu_i) ||
        !valid_grid(m_i))
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !valid_pfield(m_p) || !valid_pfield(u_p) ||
                   !valid_pfield(c_p)))
        return PIC_ERR_ARG;
    if (bins && !bins->perm) return PIC_ERR_ARG;
    for (int d = 0; d < 3; ++d)
        if (mu_i->lo[d] != m_i->lo[d] || mu_i->n[d] != m_i->n[d]) return PIC_ERR_ARG;
    if (flags & PIC_P2G_DETERMINISTIC) return PIC_ERR_UNSUPPORTED;  // APIC only
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    const bool stag = stag_dim(entity) >= 0;
    if (flags & PIC_P2G_ZERO_FIRST) {
        int rc = zero_grid(ctx, dtype, *m_i, 1);
        if (rc) return rc;
        rc = zero_grid(ctx, dtype, *mu_i, stag ? 1 : ncomp);
        if (rc) return rc;
    }
    if (np == 0) return PIC_OK;
    if (bins && (stag || ncomp == 3)) {
        const int rc = dtype == PIC_F64
                           ? polypic_p2g_brick_run<double>(ctx, order, entity, ncomp, *mesh, np, x,
                                                           m_p, u_p, c_p, *mu_i, m_i, dt, *bins)
                           : polypic_p2g_brick_run<float>(ctx, order, entity, ncomp, *mesh, np, x,
                                                          m_p, u_p, c_p, *mu_i, m_i, dt, *bins);
        if (rc != PIC_ERR_UNSUPPORTED) return rc;
    }
    const uint32_t* perm = bins ? bins->perm : nullptr;
    TimedLaunch tl(ctx, "polypic_p2g");
    if (dtype == PIC_F64) {
        PolyLaunch<double> L{make_poly_args<double>(ctx, entity, ncomp, *mesh, np, x, m_p, u_p, c_p,
                                                    *mu_i, m_i, dt, perm),
                             ctx->stream, true};
        L.run(order, ncomp, stag);
    } else {
        PolyLaunch<float> L{make_poly_args<float>(ctx, entity, ncomp, *mesh, np, x, m_p, u_p, c_p,
                                                  *mu_i, m_i, dt, perm),
                            ctx->stream, true};
        L.run(order, ncomp, stag);
    }
    return launch_status();
}

extern "C" int pic_polypic_g2p(pic_context* ctx, int order, int entity, int dtype,
                               const pic_mesh* mesh, int64_t np, int ncomp,
                               const pic_pfield* x, const pic_grid* u_i,
                               const pic_pfield* c_p, const pic_bins* bins)
{
    Region region("picasso_amd::PolyPIC::g2p");
    if (!ctx || !mesh || !(mesh->h > 0.0) || order < 1 || order > 3 ||
        !valid_entity(entity) || (dtype != PIC_F64 && dtype != PIC_F32) || np < 0 ||
        np > int64_t(UINT32_MAX) || ncomp < 1 || ncomp > 3 || !valid_grid(u_i))
        return PIC_ERR_ARG;
    if (np > 0 && (!valid_pfield(x) || !valid_pfield(c_p))) return PIC_ERR_ARG;
    if (bins && !bins->perm) return PIC_ERR_ARG;
    if (np == 0) return PIC_OK;
    DeviceGuard g(ctx->device);
    if (!g.ok) return PIC_ERR_HIP;
    const bool stag = stag_dim(entity) >= 0;
    if (bins && (stag || ncomp == 3)) {
        const int rc = dtype == PIC_F64
                           ? polypic_g2p_brick_run<double>(ctx, order, entity, ncomp, *mesh, np, x,
                                                           c_p, *u_i, *bins)
                           : polypic_g2p_brick_run<float>(ctx, order, entity, ncomp, *mesh, np, x,
                                                          c_p, *u_i, *bins);
        if (rc != PIC_ERR_UNSUPPORTED) return rc;
    }
    const uint32_t* perm = bins ? bins->perm : nullptr;
    TimedLaunch tl(ctx, "polypic_g2p");
    if (dtype == PIC_F64) {
        PolyLaunch<double> L{make_poly_args<double>(ctx, entity, ncomp, *mesh, np, x, nullptr,
                                                    nullptr, c_p, *u_i, nullptr, 0.0, perm),
                             ctx->stream, false};
        L.run(order, ncomp, stag);
    } else {
        PolyLaunch<float> L{make_poly_args<float>(ctx, entity, ncomp, *mesh, np, x, nullptr,
                                                  nullptr, c_p, *u_i, nullptr, 0.0, perm),
                            ctx->stream, false};
        L.run(order, ncomp, stag);
    }
    return launch_status();
}
