// Device side of the APIC brick scheme for particles binned by
// pic_bin_sort: the per-brick p2g and g2p bodies that the single-instance
// kernels (pic_apic_brick.hip) and the batched multi-instance kernels
// (pic_batch.hip) run for each brick of a walk.
//
// The bins are aligned with the spline's knot base (node-centred dual cells
// for quadratic splines on nodes, mesh cells for linear/cubic), so every
// particle of a bin shares one stencil and a brick's particles touch a fixed
// (edge + order)^3 node footprint.  Workgroups walk the non-empty bricks in
// list order (pic_brick.hpp), so neighbouring bricks run at about the same
// time and share their halo nodes in L2.
//
// p2g (4^3-bin bricks): four work-items per bin, one per grid component
//   (mass, momentum x/y/z).  Each walks every particle of its bin, evaluates
//   the spline and accumulates its component at the bin's NK^3 stencil nodes
//   in registers, without any cross-lane or LDS traffic.  The bin sums are
//   then stored to LDS (plain stores, one slot per bin and stencil offset),
//   and each work-item of a node of the footprint gathers the up-to-NK^3
//   sums that land on it, adding them in FP64 (exact enough to stand in for
//   compensated FP32 summation, and an FP64 add costs one FP32 issue slot on
//   CDNA4), then adds the node total into HBM with one atomic per component.
//   Nodes shared with neighbouring bricks receive one partial per brick.
// g2p: the footprint of grid values is staged in LDS with row loads (from
//   the velocity, or from momentum / mass when normalisation is fused in);
//   each work-item then reconstructs one particle from LDS.
// A particle whose stencil is not where its bin says (positions changed
// since the sort) is appended to a list that the scatter/gather kernels
// process afterwards, so results never depend on the bins being current.
#pragma once

#include "pic_apic.hpp"
#include "pic_brick.hpp"
#include "pic_sort.hpp"

namespace pic {

constexpr int BT = 256;  // work-items per g2p brick workgroup (batches, FP64, F update)

// Work-items per workgroup and particles per work-item at a time of a
// single-instance g2p brick kernel: 256 x 3 for every variant.  Under the
// dynamic brick walk it beats 128 x 4 on the 256^3 sphere in FP32 (g2p
// 0.253 against 0.262 ms inside the one-call step) and FP64 (0.437 against
// 0.461 ms); 128 x 4 was the faster FP32 shape only while a static walk
// left a tail of slow workgroups.  The collocated kernel without the F
// update also runs 128 x 4 on request (PIC_OPT_G2P_GROUP).
template <typename T, bool GF, bool STAG>
struct G2PShape {
    static constexpr int NT = BT;
    static constexpr int R = 3;
};
constexpr int PB = 4;    // p2g brick edge in bins

template <typename T>
struct BrickArgs {
    ApicArgs<T> a;
    PIC_BRICK_WALK_FIELDS
};

// Work split of the p2g brick kernel: four work-items per bin of a 4^3
// brick, one per grid component (mass, momentum x/y/z), each walking every
// particle of its bin.
//
// LDS, two uses of one buffer:
//  - particle staging: the brick's particles are copied in chunks of CAP
//    with coalesced loads (consecutive work-items, consecutive particles).
//    Each particle carries f (the fractional knot coordinate: knot base
//    found once per particle here, in FP64, and checked against the bin; a
//    mismatch is marked with f0 = STALE_F), m, and per momentum component d
//    the row (u_d, B_d0, B_d1, B_d2).  A bin group reads (f, m) and its own
//    component's row from LDS instead of issuing eight gathers to memory.
//    FP32 keeps 16 rows of CAP scalars (f0..f2, m, then c_p in storage
//    order) filled by direct global->LDS copies for m and c_p, so every
//    load of a chunk is in flight at once without holding registers
//    (measured 0.516 -> 0.502 ms on the 256^3 sphere).  FP64 stores four
//    4-scalar groups per particle from registers, (f0, f1, f2, m) and the
//    three momentum rows, read with one wide load each and XOR-swizzled by
//    (q >> 2) & 3 so 16 consecutive particles written at once hit distinct
//    bank groups.
//  - bin sums, after the last chunk: [stencil offset][bin][component], the
//    CPR components of a round contiguous, so a footprint node reads each
//    (offset, bin) slot with one wide LDS load; staged in rounds of CPR
//    components to bound the buffer (34 KB FP32, 17 KB FP64).
template <int ORDER, typename T>
struct P2GShape {
    static constexpr int NK = ORDER + 1;
    static constexpr int NN = NK * NK * NK;
    static constexpr int FN = PB + ORDER;
    static constexpr int FV = FN * FN * FN;
    static constexpr int BINS = PB * PB * PB;
    static constexpr int NT = BINS * 4;  // work-items per workgroup
    static constexpr int SLOTS = NN * BINS;  // (offset, bin) slots
    static constexpr int SUM_BUDGET = sizeof(T) == 4 ? 34 * 1024 : 17 * 1024;
    static constexpr int NROUND = 4 * SLOTS * int(sizeof(T)) <= SUM_BUDGET   ? 1
                                  : 2 * SLOTS * int(sizeof(T)) <= SUM_BUDGET ? 2
                                                                              : 4;
    static constexpr int CPR = 4 / NROUND;  // components per round
    // Particles per staging chunk: FP32 fits a typical brick (8 ppc x 64
    // bins) in one chunk at four workgroups per compute unit (616 x 64 B of
    // staging plus 1 KB of ranges stays under 40 KB); FP64, register-limited
    // to three workgroups, stages one particle per work-item (256 x 128 B),
    // two chunks for a typical brick (measured: 200 -> 1.46 ms, 256 -> 1.26
    // ms, 320 -> 1.42 ms on the 256^3 sphere).
    static constexpr int CAP = sizeof(T) == 4 ? 616 : 256;
    // FP32 stages by direct global->LDS copies into rows of CAP scalars:
    // f0..f2 (rows 0-2, written from registers), m (row ROW_M) and the 12
    // scalars of c_p in storage order (rows ROW_C..ROW_C+11).
    static constexpr bool DMA = sizeof(T) == 4;
    // Quadratic: the staged v_d are replaced, once per particle, by
    // U_d = v_d - h D^-1 sum_a B_da f_a, which leaves the accumulation loop
    // fewer vector instructions per particle (p2g_component_pre).
    static constexpr bool PRE = ORDER == 2;
    // FP64 quadratic (staged from registers) also folds m / 8 and h D^-1
    // into the staged scalars and stages g = f + 1/2 for quad_weights2:
    // group 0 holds (g0, g1, g2, m / 8), group 1 + d holds m U_d / 8 and
    // m h D^-1 B_da / 8.
    static constexpr bool FOLD = PRE && !DMA;
    // FP32 quadratic (staged by direct copies): the position rows hold
    // g = f + 1/2 and the mass row m / 8, for quad_weights2; a particle the
    // brick skips is staged with zero mass, U_d and B_d (and g = 1/2), so the
    // accumulation loop needs no per-particle test.
    static constexpr bool G2 = PRE && DMA;
    static constexpr int ROW_M = 3, ROW_C = 4;
    static constexpr int STAGE = CAP * 16;                   // scalars
    static constexpr int SUMS = CPR * SLOTS;
    static constexpr int LDS = STAGE > SUMS ? STAGE : SUMS;
    // Waves per SIMD the register allocation is held to: LDS admits four
    // workgroups (four waves per SIMD) for FP32, and the bound keeps the
    // quadratic kernel there (106 registers without spills; unbounded the
    // compiler takes 144, which leaves room for three).
    static constexpr int WAVES =
        sizeof(T) == 8 ? (ORDER == 2 ? 3 : 1) : (ORDER == 2 ? 4 : ORDER == 3 ? 3 : 1);
};

// Accumulate one particle's component into the bin's stencil sums.  v and
// B0..B2 are row `comp - 1` of u_p and B_p for a momentum component; the
// mass component passes v = 1 and B = 0, which yields exactly the weighted
// mass with the same instruction stream.
template <int ORDER, typename T>
__device__ __forceinline__ void p2g_component(const Spline<ORDER, T>& sd, T m, T v, T B0, T B1,
                                              T B2, T Dinv, T* acc)
{
    constexpr int NK = ORDER + 1;
    if constexpr (ORDER >= 2) {
        // m w_i w_j w_k (v + D^-1 (B0 d0_i + B1 d1_j + B2 d2_k))
        // (Picasso_APIC.hpp:110-138), factored per dimension.
        const T DB0 = Dinv * B0, DB1 = Dinv * B1, DB2 = Dinv * B2;
        T ai[NK], wim[NK], bj[NK], wck[NK];
#pragma unroll
        for (int i = 0; i < NK; ++i) {
            ai[i] = v + DB0 * sd.d[0][i];
            wim[i] = sd.w[0][i] * m;
            bj[i] = DB1 * sd.d[1][i];
            wck[i] = sd.w[2][i] * (DB2 * sd.d[2][i]);
        }
#pragma unroll
        for (int i = 0; i < NK; ++i)
#pragma unroll
            for (int j = 0; j < NK; ++j) {
                const T P = wim[i] * sd.w[1][j];
                const T Tij = ai[i] + bj[j];
#pragma unroll
                for (int k = 0; k < NK; ++k)
                    acc[(i * NK + j) * NK + k] += P * (sd.w[2][k] * Tij + wck[k]);
            }
    } else {
        // Gradient form m (w v + B grad w) (Picasso_APIC.hpp:287-318),
        // factored per dimension.
        T al[NK], be[NK];
#pragma unroll
        for (int i = 0; i < NK; ++i) {
            be[i] = m * sd.w[0][i];
            al[i] = m * (v * sd.w[0][i] + B0 * sd.g[0][i]);
        }
#pragma unroll
        for (int i = 0; i < NK; ++i)
#pragma unroll
            for (int j = 0; j < NK; ++j) {
                const T gam = al[i] * sd.w[1][j] + be[i] * (B1 * sd.g[1][j]);
                const T del = be[i] * sd.w[1][j] * B2;
#pragma unroll
                for (int k = 0; k < NK; ++k)
                    acc[(i * NK + j) * NK + k] += sd.w[2][k] * gam + sd.g[2][k] * del;
            }
    }
}

// p2g_component of a quadratic spline with wm = m w_i(f0), w1 = w_j(f1),
// w2 = w_k(f2), and U = v - h D^-1 (B0 f0 + B1 f1 + B2 f2) formed at staging.
// With d_a = h (o_a - f_a), o = (i, j, k) - 1, the APIC sum
// m w_i w_j w_k (v + D^-1 (B0 d0 + B1 d1 + B2 d2)) (Picasso_APIC.hpp:110-138)
// is m w_i w_j w_k (U + h D^-1 (B0 o_0 + B1 o_1 + B2 o_2)).
// Twice the quadratic B-spline weights from g = f + 1/2 in [0, 1):
// w_0 = (1 - g)^2 / 2, w_2 = g^2 / 2 and w_1 = 1 - w_0 - w_2.
template <typename T>
__device__ __forceinline__ void quad_weights2(T g, T W[3])
{
    const T a = T(1) - g;
    W[0] = a * a;
    W[2] = g * g;
    W[1] = (T(2) - W[0]) - W[2];
}

// With SCALED, B0..B2 arrive already multiplied by h D^-1 (and wm, U, B by
// any common factor).
template <bool SCALED = false, typename T>
__device__ __forceinline__ void p2g_component_pre(const T wm[3], const T w1[3], const T w2[3],
                                                  T U, T B0, T B1, T B2, T hDinv, T* acc)
{
    const T e0 = SCALED ? B0 : hDinv * B0, e1 = SCALED ? B1 : hDinv * B1,
            e2 = SCALED ? B2 : hDinv * B2;
    const T ai[3] = {U - e0, U, U + e0};
    const T wck0 = -(w2[0] * e2), wck2 = w2[2] * e2;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const T P = wm[i] * w1[j];
            const T Tij = j == 0 ? ai[i] - e1 : (j == 1 ? ai[i] : ai[i] + e1);
            T* a = acc + (i * 3 + j) * 3;
            a[0] += P * (w2[0] * Tij + wck0);
            a[1] += P * (w2[1] * Tij);
            a[2] += P * (w2[2] * Tij + wck2);
        }
}

template <int ORDER, typename T>
using P2GState = BrickStage<(P2GShape<ORDER, T>::CAP + P2GShape<ORDER, T>::NT - 1) /
                                P2GShape<ORDER, T>::NT,
                            P2GShape<ORDER, T>::NT, P2GShape<ORDER, T>::BINS>;

// PLAIN flushes with plain read-modify-writes (walks over bricks with
// disjoint footprints); DYN is the walk's kind.
template <int ORDER, typename T, bool STAG, bool PLAIN = false, bool DYN = true>
__device__ __forceinline__ void p2g_one_brick(const BrickArgs<T>& g, const BrickItem& it,
                                              BrickWalk<DYN, BrickArgs<T>>& w, T* lds,
                                              P2GState<ORDER, T>& st)
{
    using S = P2GShape<ORDER, T>;
    using T4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    constexpr int NK = S::NK, NN = S::NN, BINS = S::BINS;
    constexpr int NT = S::NT, CAP = S::CAP;
    constexpr int K = (CAP + NT - 1) / NT;
    const ApicArgs<T>& a = g.a;
    const uint32_t beg = it.beg, end = it.end;
    int f0[3];
    brick_origin<PB>(g, it.b, f0);
    const uint32_t kb = it.b * BINS;

    // Work-item group l (four work-items, one per component) accumulates
    // bin l of the brick.
    const int l = int(threadIdx.x) >> 2, comp = int(threadIdx.x) & 3;
    // Staggered: every momentum work-item works on component sdim (the
    // one grid component), and only mass and that sum are written.
    const int grp = STAG ? 1 + a.sdim : (comp > 0 ? comp : 1);
    const T hh = (T)a.lat.h;
    const T Dinv = T(ORDER == 3 ? 3.0 : 4.0) / (hh * hh);  // APIC::inertialScaling (:33-47)
    const T hDinv = hh * Dinv;
    // Lane constants of the G2 loop: the momentum work-items use U and
    // h D^-1 B; the mass work-item U * 0 + 1 and 0 * B.
    const T us = comp > 0 ? T(1) : T(0), ut = T(1) - us, hDl = us * hDinv;
    T acc[NN];
#pragma unroll
    for (int n = 0; n < NN; ++n) acc[n] = T(0);
    for (uint32_t c0 = beg; c0 < end; c0 += CAP) {
        const uint32_t c1 = min(end, c0 + uint32_t(CAP));
        uint32_t* rng = st.begin_chunk(g, a.perm, beg, c0, c1);
        if constexpr (S::DMA) {
            // Positions into registers, m and B_p straight into their LDS
            // rows: every load of the chunk is in flight before any result
            // is used, with no registers held for the 13 copied scalars.
            const int wq = int(threadIdx.x) & ~63;  // the wave's first position
            // The slot indices come either from the prefetch or from loads
            // just issued; one wait here lets the compiler count the loads
            // below exactly instead of waiting before each slot's batch.
            __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
            T xs[K][3];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (c0 + threadIdx.x + uint32_t(k * NT) >= c1) continue;
                const int64_t xb = a.x.base(int64_t(st.cs.perm[k]));
#pragma unroll
                for (int dim = 0; dim < 3; ++dim) xs[k][dim] = a.x.at(xb, dim);
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (c0 + threadIdx.x + uint32_t(k * NT) >= c1) continue;
                const int64_t p = int64_t(st.cs.perm[k]);
                const int64_t cb = a.c.base(p);
                T* row = lds + (wq + k * NT);
                load_to_lds(&a.m.at(a.m.base(p), 0), row + S::ROW_M * CAP);
#pragma unroll
                for (int e = 0; e < 12; ++e) load_to_lds(&a.c.at(cb, e), row + (S::ROW_C + e) * CAP);
            }
            T frk[K][3];
            bool skip[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                skip[k] = false;
                const uint32_t s = c0 + threadIdx.x + uint32_t(k * NT);
                if (s >= c1) continue;
                const int q = int(s - c0);
                const int lb = int(st.cs.key[k] - kb);
                const int eb[3] = {f0[0] + lb / (PB * PB), f0[1] + (lb / PB) % PB, f0[2] + lb % PB};
                T* fr = frk[k];
                bool ok = true;
                T gr[3];
#pragma unroll
                for (int dim = 0; dim < 3; ++dim) {
                    double fd;
                    const double kb0 = knot_base<ORDER>(double(xs[k][dim]), a.lat.o[dim],
                                                         a.lat.rdx, fd);
                    ok = ok && int(kb0) + Knots<ORDER>::FIRST == eb[dim];
                    fr[dim] = T(fd);
                    gr[dim] = S::G2 ? T(fd + 0.5) : T(fd);
                }
                if (!ok) {
                    push_stale(g, s);
                    skip[k] = true;
                    if constexpr (S::G2) {
#pragma unroll
                        for (int dim = 0; dim < 3; ++dim) {
                            fr[dim] = T(0);
                            gr[dim] = T(0.5);
                        }
                    } else {
                        gr[0] = T(STALE_F);
                    }
                }
#pragma unroll
                for (int dim = 0; dim < 3; ++dim) lds[dim * CAP + q] = gr[dim];
                st.cs.template note<BINS>(k, q, kb, rng);
            }
            __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
            if constexpr (S::PRE) {
                // U_d over v_d and m / 8 over m; this work-item's own copies
                // have landed.
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const uint32_t s = c0 + threadIdx.x + uint32_t(k * NT);
                    if (s >= c1) continue;
                    T* cr = lds + S::ROW_C * CAP + int(s - c0);
                    T* mr = lds + S::ROW_M * CAP + int(s - c0);
                    if (skip[k]) {
                        *mr = T(0);
#pragma unroll
                        for (int e = 0; e < 12; ++e) cr[e * CAP] = T(0);
                        continue;
                    }
                    const T* fr = frk[k];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const T sB = cr[(3 + d) * CAP] * fr[0] + cr[(6 + d) * CAP] * fr[1] +
                                     cr[(9 + d) * CAP] * fr[2];
                        cr[d * CAP] = cr[d * CAP] - hDinv * sB;
                    }
                    *mr = T(0.125) * *mr;
                }
            }
        } else {
        // Stage the chunk: consecutive work-items, consecutive particles.
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t s = c0 + threadIdx.x + uint32_t(k * NT);
            if (s >= c1) continue;
            const int64_t p = int64_t(st.cs.perm[k]);
            const int64_t xb = a.x.base(p), cb = a.c.base(p);
            const int q = int(s - c0);
            const int lb = int(st.cs.key[k] - kb);
            const int eb[3] = {f0[0] + lb / (PB * PB), f0[1] + (lb / PB) % PB, f0[2] + lb % PB};
            T4 v;
            T* fr = &v.x;
            T f3[3];
            bool ok = true;
#pragma unroll
            for (int dim = 0; dim < 3; ++dim) {
                double fd;
                const double kb0 = knot_base<ORDER>(double(a.x.at(xb, dim)), a.lat.o[dim],
                                                     a.lat.rdx, fd);
                ok = ok && int(kb0) + Knots<ORDER>::FIRST == eb[dim];
                f3[dim] = T(fd);
                fr[dim] = S::FOLD ? T(fd + 0.5) : T(fd);
            }
            if (!ok) {
                push_stale(g, s);
                v.x = T(STALE_F);
            }
            const T m8 = S::FOLD ? T(0.125) * a.m.at(a.m.base(p), 0) : T(1);
            v.w = S::FOLD ? m8 : a.m.at(a.m.base(p), 0);
            *reinterpret_cast<T4*>(lds + stage_at(q, 0)) = v;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                // B(d, a) = c_p(1 + a, d) (Picasso_APIC.hpp:94-101).
                v.x = a.c.at(cb, d);
                v.y = a.c.at(cb, 3 + d);
                v.z = a.c.at(cb, 6 + d);
                v.w = a.c.at(cb, 9 + d);
                if constexpr (S::PRE) v.x -= hDinv * (v.y * f3[0] + v.z * f3[1] + v.w * f3[2]);
                if constexpr (S::FOLD) {
                    const T mh = m8 * hDinv;
                    v.x *= m8;
                    v.y *= mh;
                    v.z *= mh;
                    v.w *= mh;
                }
                *reinterpret_cast<T4*>(lds + stage_at(q, 1 + d)) = v;
            }
            st.cs.template note<BINS>(k, q, kb, rng);
        }
        }
        st.end_chunk();
        __syncthreads();
        if (c0 == beg) w.fetch();
        const uint32_t q0 = rng[l], q1 = rng[BINS + l];
        for (uint32_t q = q0; q < q1; ++q) {
            if constexpr (S::G2) {
                // Rows: g0 g1 g2, m / 8, then U_d = row d and
                // B(d, a) = c(1 + a, d) = row 3 + 3a + d of the c block.
                // The mass work-item takes U = 1, B = 0 through its lane
                // constants (us, ut, hDl).
                const T g0 = lds[q], g1 = lds[CAP + q], g2 = lds[2 * CAP + q];
                const T m8 = lds[S::ROW_M * CAP + q];
                const T* cr = lds + (S::ROW_C + grp - 1) * CAP + q;
                const T U = cr[0], B0 = cr[3 * CAP], B1 = cr[6 * CAP], B2 = cr[9 * CAP];
                T wm[3], w1[3], w2[3];
                quad_weights2(g0, wm);
                quad_weights2(g1, w1);
                quad_weights2(g2, w2);
#pragma unroll
                for (int i = 0; i < 3; ++i) wm[i] *= m8;
                p2g_component_pre(wm, w1, w2, U * us + ut, B0, B1, B2, hDl, acc);
                continue;
            }
            if constexpr (S::PRE) {
                T4 fm, ub;
                if constexpr (S::DMA) {
                    fm.x = lds[q];
                    if (fm.x == T(STALE_F)) continue;
                    fm.y = lds[CAP + q];
                    fm.z = lds[2 * CAP + q];
                    fm.w = lds[S::ROW_M * CAP + q];
                    // U_d = row d, B(d, a) = c(1 + a, d) = row 3 + 3a + d
                    // of the c block.
                    const T* cr = lds + (S::ROW_C + grp - 1) * CAP + q;
                    ub.x = cr[0];
                    ub.y = cr[3 * CAP];
                    ub.z = cr[6 * CAP];
                    ub.w = cr[9 * CAP];
                } else {
                    fm = *reinterpret_cast<const T4*>(lds + stage_at(int(q), 0));
                    if (fm.x == T(STALE_F)) continue;
                    ub = *reinterpret_cast<const T4*>(lds + stage_at(int(q), grp));
                }
                T wm[3], w1[3], w2[3];
                if constexpr (S::FOLD) {
                    quad_weights2(fm.x, wm);
                    quad_weights2(fm.y, w1);
                    quad_weights2(fm.z, w2);
                    // The mass work-item uses U = m / 8, B = 0.
                    if (comp == 0) {
                        ub.x = fm.w;
                        ub.y = ub.z = ub.w = T(0);
                    }
                    p2g_component_pre<true>(wm, w1, w2, ub.x, ub.y, ub.z, ub.w, hDinv, acc);
                    continue;
                }
                bspline_w<2, T>(fm.x, wm);
                bspline_w<2, T>(fm.y, w1);
                bspline_w<2, T>(fm.z, w2);
#pragma unroll
                for (int i = 0; i < 3; ++i) wm[i] *= fm.w;
                // The mass work-item uses U = 1, B = 0.
                if (comp == 0) {
                    ub.x = T(1);
                    ub.y = ub.z = ub.w = T(0);
                }
                p2g_component_pre(wm, w1, w2, ub.x, ub.y, ub.z, ub.w, hDinv, acc);
                continue;
            }
            T4 fm, ub;
            if constexpr (S::DMA) {
                fm.x = lds[q];
                if (fm.x == T(STALE_F)) continue;
                fm.y = lds[CAP + q];
                fm.z = lds[2 * CAP + q];
                fm.w = lds[S::ROW_M * CAP + q];
                // u_d = c(0, d), B(d, a) = c(1 + a, d): rows d, 3 + d, 6 + d,
                // 9 + d of the c block.
                const T* cr = lds + (S::ROW_C + grp - 1) * CAP + q;
                ub.x = cr[0];
                ub.y = cr[3 * CAP];
                ub.z = cr[6 * CAP];
                ub.w = cr[9 * CAP];
            } else {
                fm = *reinterpret_cast<const T4*>(lds + stage_at(int(q), 0));
                if (fm.x == T(STALE_F)) continue;
                ub = *reinterpret_cast<const T4*>(lds + stage_at(int(q), grp));
            }
            if (comp == 0) {
                ub.x = T(1);
                ub.y = ub.z = ub.w = T(0);
            }
            const T fr[3] = {fm.x, fm.y, fm.z};
            Spline<ORDER, T> sd;
            spline_from_frac<ORDER, T, ORDER == 1, ORDER != 1>(a.lat, fr, sd);
            p2g_component<ORDER, T>(sd, fm.w, ub.x, ub.y, ub.z, ub.w, Dinv, acc);
        }
        __syncthreads();
    }

    for (int r = 0; r < S::NROUND; ++r) {
        const int r_lo = r * S::CPR;
        if (comp >= r_lo && comp < r_lo + S::CPR) {
            T* dst = lds + l * S::CPR + (comp - r_lo);
#pragma unroll
            for (int n = 0; n < NN; ++n) dst[n * BINS * S::CPR] = acc[n];
        }
        if (r == 0) w.publish();
        __syncthreads();
        if (r == 0) st.prefetch_next(g, a.perm, w, CAP);
        gather_flush<NK, PB, NT, S::CPR, T, PLAIN>(lds, f0, r_lo, 1, a.gm, a.gmu, a.err,
                                                   STAG ? 2 : 4);
        if (r + 1 < S::NROUND) __syncthreads();
    }
}

template <int BB, int ORDER, typename T, bool GF, bool STAG, int NT = BT, int R = 3,
          bool DYN = false>
__device__ __forceinline__ void g2p_one_brick(const BrickArgs<T>& g, const BrickItem& it,
                                              BrickWalk<DYN, BrickArgs<T>>& w, T* us)
{
    constexpr int FN = BB + ORDER;
    constexpr int FV = FN * FN * FN;
    const uint32_t beg = it.beg, end = it.end;
    int f0[3];
    brick_origin<BB>(g, it.b, f0);
    const ApicArgs<T>& a = g.a;
    const GView<T>& gu = a.gmu;
    w.fetch();
    for (int q = threadIdx.x; q < FV; q += NT) {
        const int ni = f0[0] + q / (FN * FN), nj = f0[1] + (q / FN) % FN, nk = f0[2] + q % FN;
        const bool in = gu.covers(0, ni, ni) && gu.covers(1, nj, nj) && gu.covers(2, nk, nk);
        const int64_t off = in ? gu.off(ni, nj, nk) : 0;
        const int64_t moff = in && a.gnm.data ? a.gnm.off(ni, nj, nk) : 0;
        T u[3] = {T(0), T(0), T(0)};
        if constexpr (STAG) {
            // One scalar grid (the staggered component).
            if (in) {
                u[0] = gu.data[off];
                if (a.gnm.data) {
                    const T mm = a.gnm.data[moff];
                    u[0] = T(double(u[0]) * (mm > T(0) ? 1.0 / double(mm) : 0.0));
                }
            }
            us[q] = u[0];
        } else {
            if (in) node_velocity3(gu, a.gnm, off, moff, u);
#pragma unroll
            for (int d = 0; d < 3; ++d) us[d * FV + q] = u[d];
        }
    }
    w.publish();
    __syncthreads();

    // Up to R particles per work-item at a time, with all their position
    // loads issued before any result is stored.  The positions are first
    // loaded from the slot itself, beside the permutation entry, as if the
    // particles were stored in bin order (the usual case, where perm[s] = s):
    // then no position load waits for the permutation.  Slots whose entry
    // differs load the position again from the particle it names.
    for (uint32_t s0 = beg + threadIdx.x; s0 < end; s0 += R * NT) {
        int64_t pp[R];
        T xp[R][3];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t s = s0 + uint32_t(r * NT);
            if (s >= end) {
                pp[r] = 0;
                continue;
            }
            pp[r] = int64_t(a.perm[s]);
            const int64_t xb = a.x.base(int64_t(s));
#pragma unroll
            for (int dim = 0; dim < 3; ++dim) xp[r][dim] = a.x.at(xb, dim);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t s = s0 + uint32_t(r * NT);
            if (s >= end || pp[r] == int64_t(s)) continue;
            const int64_t xb = a.x.base(pp[r]);
#pragma unroll
            for (int dim = 0; dim < 3; ++dim) xp[r][dim] = a.x.at(xb, dim);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t s = s0 + uint32_t(r * NT);
            if (s >= end) continue;
            Spline<ORDER, T> sd;
            eval_spline<ORDER, T, GF, true>(a.lat, xp[r], sd);
            const int li = sd.s0[0] - f0[0], lj = sd.s0[1] - f0[1], lk = sd.s0[2] - f0[2];
            if (!(li >= 0 && li < BB && lj >= 0 && lj < BB && lk >= 0 && lk < BB)) {
                push_stale(g, s);
                continue;
            }
            // Footprint nodes outside the window were staged as zeros: a
            // stencil reaching them is an error, as in the per-particle
            // gather.
            if (!stencil_inside(sd, gu)) {
                atomicOr(a.err, ERRBIT_BOUNDS);
                continue;
            }
            const int lbase = (li * FN + lj) * FN + lk;
            const int64_t cb = a.c.base(pp[r]);
            if constexpr (!STAG && !GF && ORDER == 2 && std::is_same<T, float>::value) {
                // FP32 quadratic: the distances are d_n = h (n - 1 - f), so
                // sum_n w_n d_n y_n = d_1 sum_n w_n y_n + h (w_2 y_2 - w_0 y_0)
                // in each dimension.  Each level of the factored sums then
                // forms its weighted sum and its distance term from one
                // shared product (four operations per k-row instead of six),
                // and B(d, a) = d_1a v_d + h E_a at the end.  With the middle
                // node as the reference |d_1| <= h / 2, so the two terms are
                // no larger than those of the direct sum.
                const T hh = (T)a.lat.h;
                const T* w0 = sd.w[0];
                const T* w1 = sd.w[1];
                const T* w2 = sd.w[2];
#pragma unroll 1
                for (int d = 0; d < 3; ++d) {
                    const T* ud = us + d * FV + lbase;
                    T S[3], Dj[3], Tk[3];
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        T s[3], t[3];
#pragma unroll
                        for (int j = 0; j < 3; ++j) {
                            const T* u = ud + (i * FN + j) * FN;
                            const T q = w2[2] * u[2];
                            s[j] = __builtin_fmaf(w2[0], u[0], __builtin_fmaf(w2[1], u[1], q));
                            t[j] = __builtin_fmaf(-w2[0], u[0], q);
                        }
                        const T q = w1[2] * s[2];
                        S[i] = __builtin_fmaf(w1[0], s[0], __builtin_fmaf(w1[1], s[1], q));
                        Dj[i] = __builtin_fmaf(-w1[0], s[0], q);
                        Tk[i] = __builtin_fmaf(w1[0], t[0],
                                               __builtin_fmaf(w1[1], t[1], w1[2] * t[2]));
                    }
                    const T q = w0[2] * S[2];
                    const T up = __builtin_fmaf(w0[0], S[0], __builtin_fmaf(w0[1], S[1], q));
                    const T Ei = __builtin_fmaf(-w0[0], S[0], q);
                    const T Ej =
                        __builtin_fmaf(w0[0], Dj[0], __builtin_fmaf(w0[1], Dj[1], w0[2] * Dj[2]));
                    const T Ek =
                        __builtin_fmaf(w0[0], Tk[0], __builtin_fmaf(w0[1], Tk[1], w0[2] * Tk[2]));
                    // B(d, a) = c_p(1 + a, d) (Picasso_APIC.hpp:463-471)
                    a.c.at(cb, d) = up;
                    a.c.at(cb, 3 + d) = __builtin_fmaf(sd.d[0][1], up, hh * Ei);
                    a.c.at(cb, 6 + d) = __builtin_fmaf(sd.d[1][1], up, hh * Ej);
                    a.c.at(cb, 9 + d) = __builtin_fmaf(sd.d[2][1], up, hh * Ek);
                }
                continue;
            } else if constexpr (!STAG && !GF) {
                // One grid component at a time (the velocity component and
                // its B row): v_d = sum w u_d and B(d, a) = sum w u_d d_a
                // (Picasso_APIC.hpp:439-471), summed k, j, i as in
                // apic_g2p_particle, with a third of the sums live at once
                // (FP64 148 -> 122 VGPRs, four waves per SIMD: 0.508 ->
                // 0.489 ms on the 256^3 sphere; FP32 unchanged).
                constexpr int NK = ORDER + 1;
                T wd[3][NK];
#pragma unroll
                for (int e = 0; e < 3; ++e)
#pragma unroll
                    for (int n = 0; n < NK; ++n) wd[e][n] = sd.w[e][n] * sd.d[e][n];
#pragma unroll 1
                for (int d = 0; d < 3; ++d) {
                    const T* ud = us + d * FV + lbase;
                    T up = T(0), B0 = T(0), B1 = T(0), B2 = T(0);
#pragma unroll
                    for (int i = 0; i < NK; ++i) {
                        T Si = T(0), Sdi = T(0), Ti = T(0);
#pragma unroll
                        for (int j = 0; j < NK; ++j) {
                            T sij = T(0), tij = T(0);
#pragma unroll
                            for (int k = 0; k < NK; ++k) {
                                const T ue = ud[(i * FN + j) * FN + k];
                                sij += sd.w[2][k] * ue;
                                tij += wd[2][k] * ue;
                            }
                            Si += sd.w[1][j] * sij;
                            Sdi += wd[1][j] * sij;
                            Ti += sd.w[1][j] * tij;
                        }
                        up += sd.w[0][i] * Si;
                        B0 += wd[0][i] * Si;
                        B1 += sd.w[0][i] * Sdi;
                        B2 += sd.w[0][i] * Ti;
                    }
                    // B(d, a) = c_p(1 + a, d) (Picasso_APIC.hpp:463-471)
                    a.c.at(cb, d) = up;
                    a.c.at(cb, 3 + d) = B0;
                    a.c.at(cb, 6 + d) = B1;
                    a.c.at(cb, 9 + d) = B2;
                }
                continue;
            }
            T cp[12], L[9];
            apic_g2p_particle<ORDER, STAG, GF>(a, sd, [&](int i, int j, int k, T* ue) {
                const int q = lbase + (i * FN + j) * FN + k;
                ue[0] = us[q];
                if constexpr (!STAG) {
                    ue[1] = us[FV + q];
                    ue[2] = us[2 * FV + q];
                }
            }, cp, L);
#pragma unroll
            for (int e = 0; e < 12; ++e) a.c.at(cb, e) = cp[e];
            if constexpr (GF) store_F_update(a, pp[r], L);
        }
    }
}

}  // namespace pic
