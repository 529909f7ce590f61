// Shared device-side types and the fused B-spline evaluation for the gfx950
// particle<->grid transfer kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <type_traits>

#include "picasso_amd.h"

namespace pic {

// Bits of the device-side error word latched in the context.
// ERRBIT_CAPACITY: a deterministic p2g had more stale particles than it can
// order (its result is complete but not reproducible).
enum : int { ERRBIT_BOUNDS = 1, ERRBIT_CELL = 2, ERRBIT_CAPACITY = 4 };

// ---------------------------------------------------------------------------
// Particle field: element (p, c) at data[(p >> L) * tile + (p & mask) + c * cs].
template <typename T>
struct PField {
    T* data;
    int32_t L;
    int64_t tile;
    int64_t cs;

    __device__ __forceinline__ int64_t base(int64_t p) const
    {
        return (p >> L) * tile + (p & ((int64_t(1) << L) - 1));
    }
    __device__ __forceinline__ T& at(int64_t b, int c) const
    {
        return data[b + c * cs];
    }
};

template <typename T>
inline PField<T> make_pfield(const pic_pfield& f)
{
    return PField<T>{static_cast<T*>(f.data), f.vlen_log2, f.tile_stride,
                     f.cstride};
}

// Grid window: element (i,j,k,c) at data[(i-lo0)*s0 + (j-lo1)*s1 + (k-lo2)*s2
// + c*cs].
template <typename T>
struct GView {
    T* data;
    int32_t lo[3];
    int32_t n[3];
    int64_t s[3];
    int64_t cs;

    __device__ __forceinline__ int64_t off(int i, int j, int k) const
    {
        return int64_t(i - lo[0]) * s[0] + int64_t(j - lo[1]) * s[1] +
               int64_t(k - lo[2]) * s[2];
    }
    __device__ __forceinline__ bool covers(int d, int first, int last) const
    {
        return first >= lo[d] && last < lo[d] + n[d];
    }
};

template <typename T>
inline GView<T> make_gview(const pic_grid& g)
{
    GView<T> v;
    v.data = static_cast<T*>(g.data);
    for (int d = 0; d < 3; ++d) {
        v.lo[d] = g.lo[d];
        v.n[d] = g.n[d];
        v.s[d] = g.stride[d];
    }
    v.cs = g.cstride;
    return v;
}

// Entity lattice: origin of local index 0 and spacing.
struct Lattice {
    double o[3];
    double h;
    double rdx;
};

inline bool cell_centred(int entity, int d)
{
    switch (entity) {
    case PIC_NODE: return false;
    case PIC_CELL: return true;
    case PIC_FACE_I:
    case PIC_FACE_J:
    case PIC_FACE_K: return d != entity - PIC_FACE_I;
    case PIC_EDGE_I:
    case PIC_EDGE_J:
    case PIC_EDGE_K: return d == entity - PIC_EDGE_I;
    default: return false;
    }
}

// Component index of a staggered entity (entity_type::dim), -1 if collocated.
inline int stag_dim(int entity)
{
    if (entity >= PIC_FACE_I && entity <= PIC_FACE_K) return entity - PIC_FACE_I;
    if (entity >= PIC_EDGE_I && entity <= PIC_EDGE_K) return entity - PIC_EDGE_I;
    return -1;
}

inline Lattice make_lattice(const pic_mesh& m, int entity)
{
    Lattice l;
    for (int d = 0; d < 3; ++d)
        l.o[d] = m.low[d] + (cell_centred(entity, d) ? 0.5 * m.h : 0.0);
    l.h = m.h;
    l.rdx = 1.0 / m.h;
    return l;
}

// ---------------------------------------------------------------------------
// B-splines (Cabana evaluateSpline as called by createSpline,
// Picasso_ParticleInterpolation.hpp:88-105; definitions from
// doc/apic/apic_3d_{linear,quadratic,cubic}_unit_test.nb).
//   order 1: base = floor(t),     knots base..base+1,   w = (1-f, f)
//   order 2: base = floor(t+1/2), knots base-1..base+1, w = (1/2(1/2-f)^2,
//            3/4-f^2, 1/2(1/2+f)^2)
//   order 3: base = floor(t),     knots base-1..base+2, w = ((1-f)^3/6,
//            f^3/2-f^2+2/3, (1-f)^3/2-(1-f)^2+2/3, f^3/6)
// with t = (x - o)/h the logical position and f = t - base.

template <int ORDER>
struct Knots {
    static constexpr int NK = ORDER + 1;
    static constexpr int FIRST = ORDER == 1 ? 0 : -1;
};

template <int ORDER, typename T>
__device__ __forceinline__ void bspline_w(T f, T* w)
{
    if constexpr (ORDER == 1) {
        w[0] = T(1) - f;
        w[1] = f;
    } else if constexpr (ORDER == 2) {
        const T a = T(0.5) - f, b = T(0.5) + f;
        w[0] = T(0.5) * a * a;
        w[1] = T(0.75) - f * f;
        w[2] = T(0.5) * b * b;
    } else {
        const T q = T(1) - f;
        w[0] = q * q * q * T(1.0 / 6.0);
        w[1] = T(0.5) * f * f * f - f * f + T(2.0 / 3.0);
        w[2] = T(0.5) * q * q * q - q * q + T(2.0 / 3.0);
        w[3] = f * f * f * T(1.0 / 6.0);
    }
}

// d w / d t (logical units).
template <int ORDER, typename T>
__device__ __forceinline__ void bspline_g(T f, T* g)
{
    if constexpr (ORDER == 1) {
        g[0] = T(-1);
        g[1] = T(1);
    } else if constexpr (ORDER == 2) {
        g[0] = -(T(0.5) - f);
        g[1] = T(-2) * f;
        g[2] = T(0.5) + f;
    } else {
        const T q = T(1) - f;
        g[0] = T(-0.5) * q * q;
        g[1] = (T(1.5) * f - T(2)) * f;
        g[2] = (T(2) - T(1.5) * q) * q;
        g[3] = T(0.5) * f * f;
    }
}

// d^2 w / d t^2 (logical units).
template <int ORDER, typename T>
__device__ __forceinline__ void bspline_g2(T f, T* g2)
{
    if constexpr (ORDER == 1) {
        g2[0] = T(0);
        g2[1] = T(0);
    } else if constexpr (ORDER == 2) {
        g2[0] = T(1);
        g2[1] = T(-2);
        g2[2] = T(1);
    } else {
        g2[0] = T(1) - f;
        g2[1] = T(3) * f - T(2);
        g2[2] = T(1) - T(3) * f;
        g2[3] = f;
    }
}

// Logical position and knot base for one dimension.  The mapping runs in
// FP64 for both precisions: at 256 cells per unit length an FP32 position has
// only ~16 bits below the cell size, and FP64 keeps the cell index bit-exact
// with the oracle and the fraction accurate to FP32 rounding.
// Contraction is off so the products and sums round exactly as in the oracle
// and in the bin keys: cell indices are then bit-identical everywhere.
template <int ORDER>
__device__ __forceinline__ double knot_base(double x, double o, double rdx,
                                            double& f)
{
#pragma clang fp contract(off)
    const double t = (x - o) * rdx;
    const double b = ORDER == 2 ? floor(t + 0.5) : floor(t);
    f = t - b;
    return b;
}

// Dimensions binned by node-centred cell, as a mask (bit d: dimension d),
// from pic_binning::dual; -1 when the value is not one of 0, 1 and
// PIC_DUAL_DIMS | mask.
inline int dual_mask(int dual)
{
    if (dual == 0) return 0;
    if (dual == 1) return 7;
    if ((dual & ~7) == PIC_DUAL_DIMS) return dual & 7;
    return -1;
}

// Bin coordinate floor((x - low)/h + dual/2) as a double (see pic_binning).
__device__ __forceinline__ double bin_coord(double x, double low, double rdx, int dual)
{
#pragma clang fp contract(off)
    const double t = (x - low) * rdx;
    return floor(dual ? t + 0.5 : t);
}

template <int ORDER, typename T>
struct Spline {
    static constexpr int NK = Knots<ORDER>::NK;
    int s0[3];      // first knot index per dimension
    T w[3][NK];     // weights
    T g[3][NK];     // physical gradients (only when requested)
    T d[3][NK];     // physical distances x_node - x_p (only when requested)
    T g2[3][NK];    // physical second derivatives (only when requested)
};

template <int ORDER, typename T, bool GRAD, bool DIST, bool GRAD2 = false>
__device__ __forceinline__ void eval_spline(const Lattice& lat, const T xp[3],
                                            Spline<ORDER, T>& sd)
{
    constexpr int NK = Knots<ORDER>::NK;
    constexpr int FIRST = Knots<ORDER>::FIRST;
#pragma unroll
    for (int dim = 0; dim < 3; ++dim) {
        double fd;
        const double b = knot_base<ORDER>((double)xp[dim], lat.o[dim], lat.rdx, fd);
        sd.s0[dim] = (int)b + FIRST;
        const T f = (T)fd;
        bspline_w<ORDER, T>(f, sd.w[dim]);
        if constexpr (GRAD) {
            bspline_g<ORDER, T>(f, sd.g[dim]);
            const T r = (T)lat.rdx;
#pragma unroll
            for (int n = 0; n < NK; ++n) sd.g[dim][n] *= r;
        }
        if constexpr (GRAD2) {
            bspline_g2<ORDER, T>(f, sd.g2[dim]);
            const T r2 = (T)(lat.rdx * lat.rdx);
#pragma unroll
            for (int n = 0; n < NK; ++n) sd.g2[dim][n] *= r2;
        }
        if constexpr (DIST) {
            if constexpr (std::is_same<T, double>::value) {
#pragma unroll
                for (int n = 0; n < NK; ++n)
                    sd.d[dim][n] = (lat.o[dim] + (b + double(FIRST + n)) * lat.h) -
                                   (double)xp[dim];
            } else {
                const T hh = (T)lat.h;
#pragma unroll
                for (int n = 0; n < NK; ++n) sd.d[dim][n] = hh * (T(FIRST + n) - f);
            }
        }
    }
}

// Spline weights (and physical gradients / distances) from the fractional
// knot coordinate f = t - base of each dimension, as eval_spline forms them
// (the FP64 distances come from f here, not from the absolute positions).
template <int ORDER, typename T, bool GRAD, bool DIST>
__device__ __forceinline__ void spline_from_frac(const Lattice& lat, const T f[3],
                                                 Spline<ORDER, T>& sd)
{
    constexpr int NK = Knots<ORDER>::NK;
    constexpr int FIRST = Knots<ORDER>::FIRST;
#pragma unroll
    for (int dim = 0; dim < 3; ++dim) {
        bspline_w<ORDER, T>(f[dim], sd.w[dim]);
        if constexpr (GRAD) {
            bspline_g<ORDER, T>(f[dim], sd.g[dim]);
            const T r = (T)lat.rdx;
#pragma unroll
            for (int n = 0; n < NK; ++n) sd.g[dim][n] *= r;
        }
        if constexpr (DIST) {
            const T hh = (T)lat.h;
#pragma unroll
            for (int n = 0; n < NK; ++n) sd.d[dim][n] = hh * (T(FIRST + n) - f[dim]);
        }
    }
}

// Is the whole stencil inside the grid window?
template <int ORDER, typename T, typename G>
__device__ __forceinline__ bool stencil_inside(const Spline<ORDER, T>& sd,
                                               const GView<G>& g)
{
    constexpr int NK = Knots<ORDER>::NK;
    return g.covers(0, sd.s0[0], sd.s0[0] + NK - 1) &&
           g.covers(1, sd.s0[1], sd.s0[1] + NK - 1) &&
           g.covers(2, sd.s0[2], sd.s0[2] + NK - 1);
}

// Sum over the four lanes of a quad (lanes 4q..4q+3), result in all four.
// DPP quad permutations keep this on the VALU (a generic shuffle would go
// through the LDS crossbar).
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

__device__ __forceinline__ float quad_sum(float v)
{
    v += __int_as_float(dpp_i32<0xB1>(__float_as_int(v)));  // quad_perm [1,0,3,2]
    v += __int_as_float(dpp_i32<0x4E>(__float_as_int(v)));  // quad_perm [2,3,0,1]
    return v;
}

__device__ __forceinline__ double quad_sum(double v)
{
    auto perm = [](double x, auto ctrl) {
        const long long b = __double_as_longlong(x);
        const int lo = dpp_i32<decltype(ctrl)::value>(int(b & 0xffffffffll));
        const int hi = dpp_i32<decltype(ctrl)::value>(int(b >> 32));
        return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
    };
    v += perm(v, std::integral_constant<int, 0xB1>{});
    v += perm(v, std::integral_constant<int, 0x4E>{});
    return v;
}

// Hardware floating-point atomic add (global_atomic_add_f32/_f64 on gfx950).
__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// ---------------------------------------------------------------------------
// 3x3 linear algebra in the reference's operation order
// (Picasso_BatchedLinearAlgebra.hpp:4335-4350 det, :4624-4685 inverse).
template <typename T>
__device__ __forceinline__ T det3(const T* a)
{
    return a[0] * a[4] * a[8] + a[1] * a[5] * a[6] + a[2] * a[3] * a[7] -
           a[2] * a[4] * a[6] - a[1] * a[3] * a[8] - a[0] * a[5] * a[7];
}

template <typename T>
__device__ __forceinline__ void inverse3(const T* a, T det, T* r)
{
    const T di = T(1) / det;
    r[0] = (a[4] * a[8] - a[5] * a[7]) * di;
    r[1] = (a[2] * a[7] - a[1] * a[8]) * di;
    r[2] = (a[1] * a[5] - a[2] * a[4]) * di;
    r[3] = (a[5] * a[6] - a[3] * a[8]) * di;
    r[4] = (a[0] * a[8] - a[2] * a[6]) * di;
    r[5] = (a[2] * a[3] - a[0] * a[5]) * di;
    r[6] = (a[3] * a[7] - a[4] * a[6]) * di;
    r[7] = (a[1] * a[6] - a[0] * a[7]) * di;
    r[8] = (a[0] * a[4] - a[1] * a[3]) * di;
}

// F <- (I + dt L) F, then det(F) and F^{-1}, all in FP64: the matrix
// product (Picasso_BatchedLinearAlgebra.hpp:3806-3834), operator! (:4335-4350)
// and inverse (:4624-4685).
__device__ __forceinline__ void update_F(const double L[9], double dt, double F[9], double& det,
                                         double Finv[9])
{
    double A[9], Fn[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) A[q] = (q % 4 == 0 ? 1.0 : 0.0) + dt * L[q];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            Fn[i * 3 + j] = A[i * 3 + 0] * F[0 * 3 + j] + A[i * 3 + 1] * F[1 * 3 + j] +
                            A[i * 3 + 2] * F[2 * 3 + j];
#pragma unroll
    for (int q = 0; q < 9; ++q) F[q] = Fn[q];
    det = det3(F);
    inverse3(F, det, Finv);
}

}  // namespace pic
