// Bin geometry and key computation shared by the sort and the brick kernels.
#pragma once

#include "pic_common.hpp"

namespace pic {

struct KeyGeom {
    double low[3];
    double rdx;
    int32_t cell_lo[3];
    int32_t ncell[3];
    int32_t brick[3];
    int32_t nb[3];
    uint32_t bvol;
    int32_t dual;  // bit d: dimension d bins by node-centred cell
    // log2 of the brick edges and volume when all are powers of two (the
    // usual case), so the brick / bin split is shifts and masks; else -1.
    int32_t bshift[3];
    int32_t vshift;
};

KeyGeom make_key_geom(const pic_mesh& mesh, const pic_binning& b);

// Key of particle p (0xffffffff when outside the bins); *brick receives the
// brick index.
template <typename T>
__device__ __forceinline__ uint32_t cell_key(const KeyGeom& g, const PField<T>& x, int64_t p,
                                             uint32_t* brick)
{
    const int64_t xb = x.base(p);
    int32_t rel[3];
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double c = bin_coord((double)x.at(xb, d), g.low[d], g.rdx, (g.dual >> d) & 1) -
                         (double)g.cell_lo[d];
        ok = ok && (c >= 0.0) && (c < (double)g.ncell[d]);
        rel[d] = ok ? (int32_t)c : 0;
    }
    if (!ok) {
        *brick = 0;
        return 0xffffffffu;
    }
    uint32_t bi, bj, bk, li, lj, lk;
    if (g.vshift >= 0) {
        bi = uint32_t(rel[0]) >> g.bshift[0];
        bj = uint32_t(rel[1]) >> g.bshift[1];
        bk = uint32_t(rel[2]) >> g.bshift[2];
        li = uint32_t(rel[0]) & uint32_t(g.brick[0] - 1);
        lj = uint32_t(rel[1]) & uint32_t(g.brick[1] - 1);
        lk = uint32_t(rel[2]) & uint32_t(g.brick[2] - 1);
    } else {
        bi = rel[0] / g.brick[0];
        bj = rel[1] / g.brick[1];
        bk = rel[2] / g.brick[2];
        li = rel[0] % g.brick[0];
        lj = rel[1] % g.brick[1];
        lk = rel[2] % g.brick[2];
    }
    const uint32_t b = (bi * g.nb[1] + bj) * g.nb[2] + bk;
    *brick = b;
    return b * g.bvol + (li * g.brick[1] + lj) * g.brick[2] + lk;
}

}  // namespace pic
