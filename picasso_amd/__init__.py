"""MI355X (gfx950) particle<->grid transfers with Picasso's operator interface.

The operators mirror the reference's call signatures at batch granularity:

    sd = create_spline(FieldLocation.Node, order, mesh, x)     # createSpline
    APIC.p2g(ctx, m_p, c_p, m_i, mu_i, sd)                      # APIC::p2g
    APIC.g2p(ctx, u_i, c_p, sd)                                 # APIC::g2p
    PolyPIC.p2g(ctx, m_p, u_p, c_p, mu_i, m_i, dt, sd)          # PolyPIC::p2g
    PolyPIC.g2p(ctx, u_i, c_p, sd)                              # PolyPIC::g2p

(Picasso_APIC.hpp:55-544, Picasso_PolyPIC.hpp:39-359,
Picasso_ParticleInterpolation.hpp:88-105).  Tensors are torch CUDA (HIP)
tensors used only as device memory; every operator runs the HIP kernels of
libpicasso_amd.so through its C ABI on torch's current stream.

Particle tensors are 2-D [np, ncomp] with any strides (a transposed
[ncomp, np] tensor is SoA).  Grid tensors are 3-D [n0, n1, n2] (scalar) or
4-D [n0, n1, n2, ncomp], any strides, covering local indices
[lo, lo + n) of the mesh's entity lattice.
"""
import ctypes
from dataclasses import dataclass, field
from typing import Optional, Sequence

from . import _abi
from ._abi import (CELL, EDGE_I, EDGE_J, EDGE_K, F32, F64, FACE_I, FACE_J, FACE_K, NODE,
                   PicError, check, lib)

__all__ = ["Context", "Mesh", "GridField", "create_spline", "APIC", "PolyPIC", "FieldLocation",
           "grid_normalize", "g2p_gradu", "update_F", "bin_sort", "bin_count", "Binning",
           "BinSort", "PicError", "DUAL_DIMS", "aligned_dual", "InitUniform", "InitRandom",
           "initialize_particles", "select", "filter_particles", "halo_scatter",
           "halo_gather", "AoSoA", "AoSoASlice", "permute", "G2P", "P2G", "Instance", "Batch", "FieldDep", "GridOperator",
           "lib", "LIB_PATH"]

LIB_PATH = _abi.LIB_PATH


class FieldLocation:
    """Entity kinds (Picasso_FieldTypes.hpp:125-168)."""
    Node = NODE
    Cell = CELL

    @staticmethod
    def Face(dim):
        return FACE_I + dim

    @staticmethod
    def Edge(dim):
        return EDGE_I + dim


def _cell_centred(entity, d):
    """Whether the entity's lattice is cell-centred in dimension d (nodes:
    never; cells: always; Face<D>: all but D; Edge<D>: only D)."""
    if entity == NODE:
        return False
    if entity == CELL:
        return True
    if FACE_I <= entity <= FACE_K:
        return d != entity - FACE_I
    return d == entity - EDGE_I


DUAL_DIMS = 0x10  # pic_binning dual = DUAL_DIMS | mask: dimension d dual when bit d is set


def aligned_dual(order, entity):
    """The pic_binning dual value whose bins are the knot-base lattice of
    splines of `order` on `entity`, so that the brick kernels apply: per
    dimension, dual (node-centred bins) exactly when the spline is quadratic
    on a node-aligned dimension or linear/cubic on a cell-centred one."""
    mask = 0
    for d in range(3):
        if (order == 2) != _cell_centred(entity, d):
            mask |= 1 << d
    return 0 if mask == 0 else 1 if mask == 7 else DUAL_DIMS | mask


def dual_dims(dual):
    """The per-dimension flags (0/1, three of them) of a pic_binning dual."""
    if dual in (0, 1):
        return (dual,) * 3
    if dual & ~7 != DUAL_DIMS:
        raise ValueError(f"dual {dual:#x} is not 0, 1 or DUAL_DIMS | mask")
    return tuple((dual >> d) & 1 for d in range(3))


def _dtype_code(t):
    import torch
    if t.dtype == torch.float64:
        return F64
    if t.dtype == torch.float32:
        return F32
    raise TypeError(f"unsupported dtype {t.dtype}")


def _require_device(t, name):
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")


class Context:
    """Owns a pic_context bound to a device; launches on torch's current
    stream of that device (re-read on every call)."""

    def __init__(self, device=0):
        self.device = device
        h = ctypes.c_void_p()
        check(lib.pic_context_create(ctypes.byref(h), device), "pic_context_create")
        self.handle = h

    def _bind_stream(self):
        import torch
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib.pic_set_stream(self.handle, ctypes.c_void_p(s)), "pic_set_stream")
        return self.handle

    def synchronize(self):
        check(lib.pic_synchronize(self.handle), "pic_synchronize")

    def set_timing(self, on):
        check(lib.pic_set_timing(self.handle, int(bool(on))), "pic_set_timing")

    def set_g2p_group(self, work_items):
        """Work-items per workgroup of the collocated APIC g2p brick kernels
        (pic_set_option PIC_OPT_G2P_GROUP): 128, 256, or 0 for the default
        (256).  Results do not depend on it."""
        check(lib.pic_set_option(self.handle, _abi.OPT_G2P_GROUP, int(work_items)),
              "pic_set_option")

    def reset_timing(self):
        check(lib.pic_reset_timing(self.handle), "pic_reset_timing")

    def kernel_time(self, name):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        check(lib.pic_kernel_time(self.handle, name.encode(), ctypes.byref(ms), ctypes.byref(n)),
              "pic_kernel_time")
        return ms.value, n.value

    def graph_begin(self):
        """Start capturing this context's calls (on torch's current stream,
        which must not be the null stream) into a step graph."""
        self._bind_stream()
        check(lib.pic_graph_begin(self.handle), "pic_graph_begin")

    def graph_end(self):
        """Finish the capture; returns a StepGraph to replay."""
        g = ctypes.c_void_p()
        check(lib.pic_graph_end(self.handle, ctypes.byref(g)), "pic_graph_end")
        return StepGraph(self, g)

    def close(self):
        """Destroy the context.  Refused (PicError) while a step graph made
        from it is alive: close the graphs first."""
        if getattr(self, "handle", None):
            check(lib.pic_context_destroy(self.handle), "pic_context_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StepGraph:
    """A captured sequence of calls (pic_graph); launch() replays it on the
    context's current stream.  While it is alive the context's scratch
    cannot grow: an eager call needing more raises PicError
    (PIC_ERR_ALLOC)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle

    def launch(self):
        check(lib.pic_graph_launch(self.ctx._bind_stream(), self.handle), "pic_graph_launch")

    def close(self):
        if getattr(self, "handle", None):
            lib.pic_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Mesh:
    """Uniform local mesh: low = physical position of local node (0,0,0)
    (ghosts included), h = cell size (Picasso_UniformMesh.hpp:84-190)."""
    low: Sequence[float]
    h: float

    def c(self):
        return _abi.Mesh((ctypes.c_double * 3)(*[float(v) for v in self.low]), float(self.h))


@dataclass
class GridField:
    """A grid window: tensor [n0,n1,n2(,ncomp)] for local indices lo..lo+n."""
    tensor: object
    lo: Sequence[int] = (0, 0, 0)

    @property
    def ncomp(self):
        return 1 if self.tensor.dim() == 3 else self.tensor.shape[3]

    def c(self):
        t = self.tensor
        _require_device(t, "grid field")
        st = t.stride()
        return _abi.Grid(t.data_ptr(), (ctypes.c_int32 * 3)(*[int(v) for v in self.lo]),
                         (ctypes.c_int32 * 3)(*t.shape[:3]), (ctypes.c_int64 * 3)(*st[:3]),
                         st[3] if t.dim() == 4 else 0)


class AoSoA:
    """A Cabana AoSoA buffer as Picasso's particle lists hold it
    (Picasso_ParticleList.hpp:90-138): tiles of V particles; inside a tile
    every member's scalars are stored component-major, V consecutive
    particles per component.  `buffer` is a 1-D device tensor of the
    scalars of all tiles, `members` the scalar count of each member in
    order (e.g. (3, 1, 12) for x, m, c_p)."""

    def __init__(self, buffer, vector_length, members, size):
        v = int(vector_length)
        if v < 1 or v & (v - 1):
            raise ValueError("vector_length must be a power of two")
        self.buffer, self.V, self.members, self.size = buffer, v, tuple(members), int(size)
        self.q = sum(self.members)
        need = -(-self.size // v) * v * self.q
        if buffer.dim() != 1 or buffer.shape[0] < need:
            raise ValueError(f"AoSoA buffer must be 1-D with at least {need} scalars")

    def slice(self, member):
        """The particle field of one member (Cabana::slice<member>)."""
        return AoSoASlice(self, sum(self.members[:member]), self.members[member])

    @staticmethod
    def from_fields(fields, vector_length):
        """Pack [np, q_f] tensors (one per member, same dtype and device)
        into a new AoSoA."""
        import torch
        np_ = fields[0].shape[0]
        v = int(vector_length)
        q = sum(f.shape[1] for f in fields)
        nt = -(-np_ // v)
        buf = torch.zeros(nt * v * q, dtype=fields[0].dtype, device=fields[0].device)
        tiles = buf.view(nt, q, v)
        o = 0
        for f in fields:
            pad = torch.zeros((nt * v, f.shape[1]), dtype=f.dtype, device=f.device)
            pad[:np_] = f
            tiles[:, o:o + f.shape[1], :] = pad.view(nt, v, f.shape[1]).transpose(1, 2)
            o += f.shape[1]
        return AoSoA(buf, v, [f.shape[1] for f in fields], np_)


class AoSoASlice:
    """One member of an AoSoA: element (p, c) at buffer[(p // V) * V q +
    (offset + c) V + p % V], i.e. vlen_log2 = log2 V, tile stride V q and
    component stride V in the pic_pfield formula."""

    def __init__(self, aosoa, offset, ncomp):
        self.aosoa, self.offset, self.ncomp = aosoa, offset, ncomp

    @property
    def shape(self):
        return (self.aosoa.size, self.ncomp)

    @property
    def dtype(self):
        return self.aosoa.buffer.dtype

    @property
    def is_cuda(self):
        return self.aosoa.buffer.is_cuda

    @property
    def device(self):
        return self.aosoa.buffer.device

    def dim(self):
        return 2

    def to_tensor(self):
        """A [size, ncomp] copy of the member (for checks)."""
        a = self.aosoa
        nt = -(-a.size // a.V)
        t = a.buffer[:nt * a.V * a.q].view(nt, a.q, a.V)[:, self.offset:self.offset + self.ncomp, :]
        return t.transpose(1, 2).reshape(nt * a.V, self.ncomp)[:a.size]

    def c(self):
        a = self.aosoa
        esz = a.buffer.element_size()
        return _abi.PField(a.buffer.data_ptr() + self.offset * a.V * esz,
                           a.V.bit_length() - 1, a.V * a.q, a.V)


def _pfield(t, ncomp, name, np_=None, dtype=None):
    """pic_pfield of a [np, ncomp] tensor (any strides) or an AoSoA slice,
    after checking its width, its row count against np_ and its dtype
    against the call's (every check raises before anything is launched)."""
    if isinstance(t, AoSoASlice):
        if t.ncomp != ncomp:
            raise ValueError(f"{name} must have {ncomp} components, got {t.ncomp}")
    else:
        if t.dim() == 1:
            t = t.unsqueeze(1)
        if t.dim() != 2 or t.shape[1] != ncomp:
            raise ValueError(f"{name} must be [np, {ncomp}], got {tuple(t.shape)}")
    if np_ is not None and t.shape[0] < np_:
        raise ValueError(f"{name} holds {t.shape[0]} particles, fewer than np = {np_}")
    if dtype is not None and _dtype_code(t) != dtype:
        raise TypeError(f"{name} is {t.dtype}; every field of the call must share x's dtype")
    _require_device(t, name)
    if isinstance(t, AoSoASlice):
        return t.c()
    st = t.stride()
    return _abi.PField(t.data_ptr(), 0, st[0], st[1])


def _grid(g, name, dtype, ncomp=None):
    """pic_grid of a GridField after checking its dtype and, when given, its
    component count."""
    if _dtype_code(g.tensor) != dtype:
        raise TypeError(f"{name} is {g.tensor.dtype}; every field of the call must share "
                        "x's dtype")
    if ncomp is not None and g.ncomp != ncomp:
        raise ValueError(f"{name} must have {ncomp} component(s), got {g.ncomp}")
    return g.c()


def _ncomp_grid(entity, collocated):
    """Components of a transfer's momentum / velocity grid: `collocated`
    on Node/Cell, 1 on a staggered Face/Edge lattice."""
    return collocated if entity in (NODE, CELL) else 1


@dataclass
class Spline:
    """Batched stand-in for the reference's SplineData: location, order,
    mesh and the particle positions it is evaluated at."""
    location: int
    order: int
    mesh: Mesh
    x: object
    bins: Optional["BinSort"] = None


def create_spline(location, order, mesh, x, bins=None):
    """createSpline(Location, InterpolationOrder<order>, local_mesh, x, ...)
    (Picasso_ParticleInterpolation.hpp:88-105).  Weights, gradients and
    distances are evaluated inside the transfer kernels."""
    if order not in (1, 2, 3):
        raise ValueError("order must be 1, 2 or 3")
    return Spline(location, order, mesh, x, bins)


def _bins_ptr(bins, np_):
    """pic_bins of a BinSort after checking that it covers np_ particles."""
    if bins is None:
        return None
    _check_bins(bins, np_)
    return ctypes.byref(bins.c())


class APIC:
    @staticmethod
    def p2g(ctx, m_p, c_p, m_i, mu_i, sd, zero_first=False, deterministic=False):
        """APIC::p2g(m_p, c_p, m_i, mu_i, sd) (Picasso_APIC.hpp:55-402).
        c_p is [np, 12] (row-major 4x3); accumulates into m_i, mu_i.
        deterministic: bit-identical sums from run to run (needs bins the
        brick kernels take; PIC_P2G_DETERMINISTIC)."""
        x = sd.x
        np_, dt = x.shape[0], _dtype_code(x)
        args = (ctypes.byref(sd.mesh.c()), np_, ctypes.byref(_pfield(x, 3, "x")),
                ctypes.byref(_pfield(m_p, 1, "m_p", np_, dt)),
                ctypes.byref(_pfield(c_p, 12, "c_p", np_, dt)),
                ctypes.byref(_grid(m_i, "m_i", dt, 1)),
                ctypes.byref(_grid(mu_i, "mu_i", dt, _ncomp_grid(sd.location, 3))),
                _bins_ptr(sd.bins, np_),
                (_abi.P2G_ZERO_FIRST if zero_first else 0) |
                (_abi.P2G_DETERMINISTIC if deterministic else 0))
        check(lib.pic_apic_p2g(ctx._bind_stream(), sd.order, sd.location, dt, *args),
              "pic_apic_p2g")

    @staticmethod
    def g2p(ctx, u_i, c_p, sd):
        """APIC::g2p(u_i, c_p, sd) (Picasso_APIC.hpp:407-544)."""
        x = sd.x
        np_, dt = x.shape[0], _dtype_code(x)
        args = (ctypes.byref(sd.mesh.c()), np_, ctypes.byref(_pfield(x, 3, "x")),
                ctypes.byref(_grid(u_i, "u_i", dt, _ncomp_grid(sd.location, 3))),
                ctypes.byref(_pfield(c_p, 12, "c_p", np_, dt)), _bins_ptr(sd.bins, np_))
        check(lib.pic_apic_g2p(ctx._bind_stream(), sd.order, sd.location, dt, *args),
              "pic_apic_g2p")

    @staticmethod
    def g2p_normalized(ctx, m_i, mu_i, c_p, sd):
        """APIC::g2p of u_i = mu_i / m_i computed on the fly (grid
        normalisation fused into g2p)."""
        x = sd.x
        np_, dt = x.shape[0], _dtype_code(x)
        args = (ctypes.byref(sd.mesh.c()), np_, ctypes.byref(_pfield(x, 3, "x")),
                ctypes.byref(_grid(m_i, "m_i", dt, 1)),
                ctypes.byref(_grid(mu_i, "mu_i", dt, _ncomp_grid(sd.location, 3))),
                ctypes.byref(_pfield(c_p, 12, "c_p", np_, dt)), _bins_ptr(sd.bins, np_))
        check(lib.pic_apic_g2p_normalized(ctx._bind_stream(), sd.order, sd.location, dt, *args),
              "pic_apic_g2p_normalized")

    @staticmethod
    def step(ctx, m_p, c_p, m_i, mu_i, sd, geom, out=None):
        """One APIC step in one call (pic_apic_step): bin_sort of sd.x by
        geom, p2g into m_i / mu_i with the grid zeroed first, and g2p of
        mu_i / m_i into c_p -- bin_sort, APIC.p2g(zero_first=True) and
        APIC.g2p_normalized with fewer launches.  Returns the bins (out, or
        new arrays), which later calls may reuse as sd.bins."""
        import torch
        x = sd.x
        np_, dt = x.shape[0], _dtype_code(x)
        nb = bin_count(geom)
        if out is None:
            i32 = dict(dtype=torch.int32, device=x.device)
            out = BinSort(torch.empty(np_, **i32), torch.empty(np_, **i32),
                          torch.empty(nb + 1, **i32), geom, torch.empty(nb + 1, **i32))
        _check_bins(out, np_, nb)
        out.geom = geom
        args = (ctypes.byref(sd.mesh.c()), np_, ctypes.byref(_pfield(x, 3, "x")),
                ctypes.byref(_pfield(m_p, 1, "m_p", np_, dt)),
                ctypes.byref(_pfield(c_p, 12, "c_p", np_, dt)),
                ctypes.byref(_grid(m_i, "m_i", dt, 1)),
                ctypes.byref(_grid(mu_i, "mu_i", dt, _ncomp_grid(sd.location, 3))),
                ctypes.byref(geom.c()), out.keys.data_ptr(), out.perm.data_ptr(),
                out.bin_start.data_ptr(),
                out.active.data_ptr() if out.active is not None else None)
        check(lib.pic_apic_step(ctx._bind_stream(), sd.order, sd.location, dt, *args),
              "pic_apic_step")
        return out

    @staticmethod
    def g2p_update_F(ctx, u_i, c_p, dt, F, det, Finv, sd, m_i=None):
        """APIC::g2p fused with the velocity gradient and the FP64
        deformation-gradient update F <- (I + dt grad u) F, det(F), F^-1.
        With m_i given, u_i is the grid momentum and the velocity mu/m is
        formed on the fly.  F [np, 9] (updated in place), det [np, 1] and
        Finv [np, 9] are float64."""
        x = sd.x
        np_, dty = x.shape[0], _dtype_code(x)
        args = (ctypes.byref(sd.mesh.c()), np_, ctypes.byref(_pfield(x, 3, "x")),
                ctypes.byref(_grid(m_i, "m_i", dty, 1)) if m_i is not None else None,
                ctypes.byref(_grid(u_i, "u_i", dty, 3)),
                ctypes.byref(_pfield(c_p, 12, "c_p", np_, dty)), float(dt),
                ctypes.byref(_pfield(F, 9, "F", np_, F64)),
                ctypes.byref(_pfield(det, 1, "det", np_, F64)),
                ctypes.byref(_pfield(Finv, 9, "Finv", np_, F64)), _bins_ptr(sd.bins, np_))
        check(lib.pic_apic_g2p_update_F(ctx._bind_stream(), sd.order, sd.location, dty, *args),
              "pic_apic_g2p_update_F")


class PolyPIC:
    @staticmethod
    def nmode(order):
        return check(lib.pic_polypic_nmode(order), "pic_polypic_nmode")

    @staticmethod
    def p2g(ctx, m_p, u_p, c_p, mu_i, m_i, dt, sd, ncomp=3, zero_first=False):
        """PolyPIC::p2g(m_p, u_p, c_p, mu_i, m_i, dt, sd)
        (Picasso_PolyPIC.hpp:39-229).  u_p is [np, nmode*3], c_p
        [np, nmode*ncomp]."""
        x = sd.x
        np_, dty = x.shape[0], _dtype_code(x)
        nm = PolyPIC.nmode(sd.order)
        args = (ctypes.byref(sd.mesh.c()), np_, ncomp, ctypes.byref(_pfield(x, 3, "x")),
                ctypes.byref(_pfield(m_p, 1, "m_p", np_, dty)),
                ctypes.byref(_pfield(u_p, nm * 3, "u_p", np_, dty)),
                ctypes.byref(_pfield(c_p, nm * ncomp, "c_p", np_, dty)),
                ctypes.byref(_grid(mu_i, "mu_i", dty, _ncomp_grid(sd.location, ncomp))),
                ctypes.byref(_grid(m_i, "m_i", dty, 1)), float(dt), _bins_ptr(sd.bins, np_),
                _abi.P2G_ZERO_FIRST if zero_first else 0)
        check(lib.pic_polypic_p2g(ctx._bind_stream(), sd.order, sd.location, dty, *args),
              "pic_polypic_p2g")

    @staticmethod
    def g2p(ctx, u_i, c_p, sd, ncomp=3):
        """PolyPIC::g2p(u_i, c_p, sd) (Picasso_PolyPIC.hpp:235-359)."""
        x = sd.x
        np_, dty = x.shape[0], _dtype_code(x)
        nm = PolyPIC.nmode(sd.order)
        args = (ctypes.byref(sd.mesh.c()), np_, ncomp, ctypes.byref(_pfield(x, 3, "x")),
                ctypes.byref(_grid(u_i, "u_i", dty, _ncomp_grid(sd.location, ncomp))),
                ctypes.byref(_pfield(c_p, nm * ncomp, "c_p", np_, dty)), _bins_ptr(sd.bins, np_))
        check(lib.pic_polypic_g2p(ctx._bind_stream(), sd.order, sd.location, dty, *args),
              "pic_polypic_g2p")


def grid_normalize(ctx, m_i, mu_i, u_i, rel_mass_eps=0.0):
    """u_i = mu_i / m_i (0 where m_i == 0)."""
    dt = _dtype_code(m_i.tensor)
    args = (mu_i.ncomp, ctypes.byref(_grid(m_i, "m_i", dt, 1)),
            ctypes.byref(_grid(mu_i, "mu_i", dt)),
            ctypes.byref(_grid(u_i, "u_i", dt, mu_i.ncomp)), float(rel_mass_eps))
    check(lib.pic_grid_normalize(ctx._bind_stream(), dt, *args), "pic_grid_normalize")


HALO_GATHER, HALO_SCATTER_SUM = 0, 1


def _halo(ctx, op, field, own_lo, own_n, periodic):
    i3 = ctypes.c_int32 * 3
    rc = lib.pic_grid_halo(ctx._bind_stream(), op, _dtype_code(field.tensor), ctypes.byref(field.c()),
                           field.ncomp, i3(*[int(v) for v in own_lo]),
                           i3(*[int(v) for v in own_n]), i3(*[int(bool(v)) for v in periodic]))
    check(rc, "pic_grid_halo")


def halo_scatter(ctx, field, own_lo, own_n, periodic=(True, True, True)):
    """Scatter-sum halo (GridOperator after a p2g, Picasso_GridOperator.hpp:
    211-215, 489-491) on one GPU: every ghost node of a periodic dimension
    adds into its owned image own_lo + (i - own_lo) mod own_n."""
    _halo(ctx, HALO_SCATTER_SUM, field, own_lo, own_n, periodic)


def halo_gather(ctx, field, own_lo, own_n, periodic=(True, True, True)):
    """Gather halo (GridOperator before an apply, Picasso_GridOperator.hpp:
    203-208, 465-466) on one GPU: every ghost node takes its owned image."""
    _halo(ctx, HALO_GATHER, field, own_lo, own_n, periodic)


def g2p_gradu(ctx, u_i, gradu, sd):
    """Velocity gradient gradu[p, i*3+j] = du_i/dx_j (G2P::gradient)."""
    x = sd.x
    np_, dt = x.shape[0], _dtype_code(x)
    args = (ctypes.byref(sd.mesh.c()), np_, ctypes.byref(_pfield(x, 3, "x")),
            ctypes.byref(_grid(u_i, "u_i", dt, 3)),
            ctypes.byref(_pfield(gradu, 9, "gradu", np_, dt)), _bins_ptr(sd.bins, np_))
    check(lib.pic_g2p_gradu(ctx._bind_stream(), sd.order, sd.location, dt, *args),
          "pic_g2p_gradu")


INTERP_VALUE, INTERP_GRADIENT, INTERP_DIVERGENCE = 0, 1, 2


def _g2p_interp(ctx, op, sd, field, ncomp, result, nout, what):
    x = sd.x
    np_, dt = x.shape[0], _dtype_code(x)
    args = (ctypes.byref(sd.mesh.c()), np_, ctypes.byref(_pfield(x, 3, "x")),
            ctypes.byref(_grid(field, "field", dt, ncomp)), ncomp,
            ctypes.byref(_pfield(result, nout, "result", np_, dt)), _bins_ptr(sd.bins, np_))
    check(lib.pic_g2p_interp(ctx._bind_stream(), op, sd.order, sd.location, dt, *args), what)


# Grid components a P2G writes, per (operation, value components).
_P2G_FIELD_NCOMP = {(0, 1): 1, (0, 2): 2, (0, 3): 3, (1, 1): 3, (2, 3): 1, (2, 9): 3}


def _p2g_interp(ctx, op, sd, value, ncomp, field, zero_first, what):
    x = sd.x
    np_, dt = x.shape[0], _dtype_code(x)
    args = (ctypes.byref(sd.mesh.c()), np_, ctypes.byref(_pfield(x, 3, "x")),
            ctypes.byref(_pfield(value, ncomp, "value", np_, dt)), ncomp,
            ctypes.byref(_grid(field, "field", dt, _P2G_FIELD_NCOMP.get((op, ncomp)))),
            _abi.P2G_ZERO_FIRST if zero_first else 0, _bins_ptr(sd.bins, np_))
    check(lib.pic_p2g_interp(ctx._bind_stream(), op, sd.order, sd.location, dt, *args), what)


class G2P:
    """Picasso::G2P (Picasso_ParticleInterpolation.hpp:110-189) for every
    particle of sd at once: result[p] from the grid field."""

    @staticmethod
    def value(ctx, sd, field, result, ncomp=None):
        """result[p, c] = sum_n w_n field(n, c) (scalar: ncomp 1, vector: 3)."""
        nc = ncomp if ncomp is not None else (1 if result.dim() == 1 else result.shape[1])
        _g2p_interp(ctx, INTERP_VALUE, sd, field, nc, result, nc, "pic_g2p_interp(value)")

    @staticmethod
    def gradient(ctx, sd, field, result, ncomp=1):
        """Scalar field (ncomp 1): result[p, d] = d field/dx_d; vector field
        (ncomp 3): result[p, 3i + j] = d field_i/dx_j."""
        _g2p_interp(ctx, INTERP_GRADIENT, sd, field, ncomp, result, 3 * ncomp,
                    "pic_g2p_interp(gradient)")

    @staticmethod
    def divergence(ctx, sd, field, result):
        """result[p] = div field (3-component field)."""
        _g2p_interp(ctx, INTERP_DIVERGENCE, sd, field, 3, result, 1,
                    "pic_g2p_interp(divergence)")


class P2G:
    """Picasso::P2G (Picasso_ParticleInterpolation.hpp:197-279) for every
    particle of sd at once, accumulated into the grid field."""

    @staticmethod
    def value(ctx, sd, value, field, zero_first=False):
        """field(n, c) += w_n value[p, c]."""
        nc = 1 if value.dim() == 1 else value.shape[1]
        _p2g_interp(ctx, INTERP_VALUE, sd, value, nc, field, zero_first, "pic_p2g_interp(value)")

    @staticmethod
    def gradient(ctx, sd, value, field, zero_first=False):
        """field(n, d) += dw_n/dx_d value[p] (scalar values, 3-component field)."""
        _p2g_interp(ctx, INTERP_GRADIENT, sd, value, 1, field, zero_first,
                    "pic_p2g_interp(gradient)")

    @staticmethod
    def divergence(ctx, sd, value, field, zero_first=False):
        """Vector values [np, 3]: field(n) += grad w_n . value[p]; tensor
        values [np, 9] (row-major): field(n, i) += sum_j dw_n/dx_j value[p, 3i + j]."""
        nc = value.shape[1]
        _p2g_interp(ctx, INTERP_DIVERGENCE, sd, value, nc, field, zero_first,
                    "pic_p2g_interp(divergence)")


def update_F(ctx, gradu, dt, F, det, Finv):
    """F <- (I + dt gradu) F, det(F), F^{-1} in FP64."""
    np_ = F.shape[0]
    args = (_dtype_code(gradu), ctypes.byref(_pfield(gradu, 9, "gradu", np_)), float(dt),
            ctypes.byref(_pfield(F, 9, "F", np_, F64)),
            ctypes.byref(_pfield(det, 1, "det", np_, F64)),
            ctypes.byref(_pfield(Finv, 9, "Finv", np_, F64)))
    check(lib.pic_update_F(ctx._bind_stream(), np_, *args), "pic_update_F")


@dataclass
class Binning:
    """Bin lattice (pic_binning): bins floor((x-low)/h + dual_d/2) in
    [cell_lo, cell_lo+ncell), grouped in bricks of `brick` bins.  dual is 0
    (mesh cells), 1 (node-centred cells in every dimension) or
    DUAL_DIMS | mask (node-centred in the dimensions whose bit is set);
    aligned_dual(order, entity) gives the one the brick kernels need."""
    cell_lo: Sequence[int]
    ncell: Sequence[int]
    brick: Sequence[int] = (8, 8, 8)
    dual: int = 0

    def c(self):
        i3 = ctypes.c_int32 * 3
        return _abi.Binning(i3(*[int(v) for v in self.cell_lo]),
                            i3(*[int(v) for v in self.ncell]),
                            i3(*[int(v) for v in self.brick]), int(self.dual))


@dataclass
class BinSort:
    """Output of bin_sort: perm (sorted slot -> particle), keys (key of each
    sorted slot), bin_start (per brick), the list of non-empty bricks
    (active, optional) and the binning."""
    perm: object
    keys: object
    bin_start: object
    geom: Binning
    active: object = None

    def c(self):
        return _abi.Bins(self.perm.data_ptr(), self.keys.data_ptr(), self.bin_start.data_ptr(),
                         self.active.data_ptr() if self.active is not None else None,
                         self.geom.c())


def permute(ctx, perm, pairs):
    """Gather particle fields: for each (src, dst) tensor pair ([n_src, q] and
    [n, q], same dtype), dst[s] = src[perm[s]] for s < n (pic_permute); perm
    holds n indices below n_src (a bin order, or a selection)."""
    if not pairs:
        return
    dt = _dtype_code(pairs[0][0])
    np_ = pairs[0][1].shape[0]
    n = len(pairs)
    srcs = (_abi.PField * n)(*[_pfield(a, a.shape[1], "src") for a, _ in pairs])
    dsts = (_abi.PField * n)(*[_pfield(b, b.shape[1], "dst") for _, b in pairs])
    nc = (ctypes.c_int * n)(*[a.shape[1] for a, _ in pairs])
    if perm.shape[0] < np_:
        raise ValueError("permute: perm shorter than the destination")
    for a, b in pairs:
        if (_dtype_code(a) != dt or _dtype_code(b) != dt or a.shape[1] != b.shape[1]
                or b.shape[0] != np_ or a.shape[0] < np_):
            raise ValueError("permute: pairs must share dtype and width, destinations the "
                             "same length, sources at least as long")
    rc = lib.pic_permute(ctx._bind_stream(), dt, np_, ctypes.c_void_p(perm.data_ptr()), n,
                         srcs, dsts, nc)
    check(rc, "pic_permute")


@dataclass
class FieldDep:
    """A gather or scatter dependency of a GridOperator: a grid window, its
    owned entities [own_lo, own_lo + own_n) and its periodic dimensions."""
    field: GridField
    own_lo: Sequence[int]
    own_n: Sequence[int]
    periodic: Sequence[bool] = (False, False, False)

    def c(self):
        i3 = ctypes.c_int32 * 3
        return _abi.FieldDep(self.field.c(), self.field.ncomp, i3(*[int(v) for v in self.own_lo]),
                             i3(*[int(v) for v in self.own_n]),
                             i3(*[int(bool(v)) for v in self.periodic]))


class GridOperator:
    """GridOperator::apply for particle operations on one GPU
    (Picasso_GridOperator.hpp:459-496, pic_grid_operator_apply): gather the
    gather fields' periodic halos, zero the scatter fields, run body(ctx)
    (the particle transfers, e.g. G2P / P2G / APIC calls), then scatter-sum
    the scatter fields' periodic halos."""

    def __init__(self, gather=(), scatter=()):
        self.gather, self.scatter = list(gather), list(scatter)

    def apply(self, ctx, body):
        deps = self.gather + self.scatter
        if not deps:
            raise ValueError("GridOperator needs at least one field")
        dt = _dtype_code(deps[0].field.tensor)
        for d in deps:
            _grid(d.field, "field", dt)
        ga = (_abi.FieldDep * max(1, len(self.gather)))(*[d.c() for d in self.gather])
        sa = (_abi.FieldDep * max(1, len(self.scatter)))(*[d.c() for d in self.scatter])
        failure = []

        def thunk(_ctx, _user):
            try:
                body(ctx)
                return 0
            except Exception as e:  # reported after the C call returns
                failure.append(e)
                return _abi.PIC_ERR_ARG

        cb = _abi.APPLY_FN(thunk)
        rc = lib.pic_grid_operator_apply(ctx._bind_stream(), dt, len(self.gather), ga,
                                         len(self.scatter), sa, cb, None)
        if failure:
            raise failure[0]
        check(rc, "pic_grid_operator_apply")


@dataclass
class Instance:
    """One independent APIC simulation of a Batch: its mesh, its particles
    [first, first + np) of the batch's particle fields, its bins (4x4x4
    bricks on the spline's knot lattice) and its grid windows (mass, and
    3-component momentum over the same window)."""
    mesh: Mesh
    first: int
    np: int
    geom: Binning
    m_i: GridField
    mu_i: GridField


class Batch:
    """Many independent APIC instances stepped together (pic_batch_*): every
    pass is one launch over all instances' bricks, and the batch owns its
    bins and work buffers, so its calls allocate nothing and a step can be
    captured in a step graph."""

    def __init__(self, ctx, order, dtype, instances):
        import torch
        self.dt = F64 if dtype == torch.float64 else F32
        self.np = sum(int(i.np) for i in instances)
        self.instances = list(instances)  # keeps the grid tensors alive
        tab = (_abi.Instance * len(instances))()
        for k, inst in enumerate(instances):
            for g, nm in ((inst.m_i, "m_i"), (inst.mu_i, "mu_i")):
                _grid(g, nm, self.dt, 1 if nm == "m_i" else 3)
            tab[k] = _abi.Instance(inst.mesh.c(), int(inst.first), int(inst.np), inst.geom.c(),
                                   inst.m_i.c(), inst.mu_i.c())
        h = ctypes.c_void_p()
        check(lib.pic_batch_create(ctx._bind_stream(), int(order), self.dt, len(instances), tab,
                                   ctypes.byref(h)), "pic_batch_create")
        self.handle = h

    def _fields(self, x, m_p, c_p):
        out = [ctypes.byref(_pfield(x, 3, "x", self.np, self.dt))]
        if m_p is not None:
            out.append(ctypes.byref(_pfield(m_p, 1, "m_p", self.np, self.dt)))
        out.append(ctypes.byref(_pfield(c_p, 12, "c_p", self.np, self.dt)))
        return out

    def bin_sort(self, ctx, x):
        xf = ctypes.byref(_pfield(x, 3, "x", self.np, self.dt))
        check(lib.pic_batch_bin_sort(ctx._bind_stream(), self.handle, xf), "pic_batch_bin_sort")

    def p2g(self, ctx, x, m_p, c_p, zero_first=True):
        f = self._fields(x, m_p, c_p)
        check(lib.pic_batch_apic_p2g(ctx._bind_stream(), self.handle, *f,
                                     _abi.P2G_ZERO_FIRST if zero_first else 0),
              "pic_batch_apic_p2g")

    def g2p_normalized(self, ctx, x, c_p):
        f = self._fields(x, None, c_p)
        check(lib.pic_batch_apic_g2p_normalized(ctx._bind_stream(), self.handle, *f),
              "pic_batch_apic_g2p_normalized")

    def step(self, ctx, x, m_p, c_p, zero_first=True):
        """Sort, p2g and normalised g2p of every instance."""
        f = self._fields(x, m_p, c_p)
        check(lib.pic_batch_apic_step(ctx._bind_stream(), self.handle, *f,
                                      _abi.P2G_ZERO_FIRST if zero_first else 0),
              "pic_batch_apic_step")

    def close(self):
        if getattr(self, "handle", None):
            lib.pic_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


InitUniform, InitRandom = 0, 1  # Picasso_ParticleInit.hpp:30-35 (tags)


def initialize_particles(ctx, init, ppc, mesh, cell_lo, ncell, dtype=None, seed=0,
                         with_volume=True):
    """initializeParticles(InitUniform / InitRandom, ...)
    (Picasso_ParticleInit.hpp:121-248, 279-400) without the creation
    functor: positions x [np, 3] (and volumes [np, 1]) of every particle of
    the cells [cell_lo, cell_lo + ncell), ppc per cell (InitRandom) or ppc^3
    on a lattice (InitUniform; ppc = particles_per_cell_dim), in the
    reference's particle order.  Apply the functor's choice with
    filter_particles."""
    import torch
    dtype = dtype or torch.float64
    n_cells = 1
    for v in ncell:
        n_cells *= int(v)
    np_ = n_cells * (ppc ** 3 if init == InitUniform else ppc)
    x = torch.empty((3, np_), dtype=dtype, device="cuda").t()
    vol = torch.empty((1, np_), dtype=dtype, device="cuda").t() if with_volume else None
    i3 = ctypes.c_int32 * 3
    rc = lib.pic_init_particles(ctx._bind_stream(), int(init), _dtype_code(x),
                                ctypes.byref(mesh.c()), i3(*[int(v) for v in cell_lo]),
                                i3(*[int(v) for v in ncell]), int(ppc), int(seed) & (2**64 - 1),
                                np_, ctypes.byref(_pfield(x, 3, "x")),
                                ctypes.byref(_pfield(vol, 1, "volume")) if vol is not None
                                else None)
    check(rc, "pic_init_particles")
    return (x, vol) if with_volume else x


def select(ctx, keep):
    """Indices of the particles a mask keeps, in increasing order, as a
    uint32-valued int32 tensor (pic_select)."""
    import torch
    _require_device(keep, "keep")
    k = keep.to(torch.uint8).contiguous()
    np_ = k.shape[0]
    kept = torch.empty(np_ + 1, dtype=torch.int32, device=k.device)
    check(lib.pic_select(ctx._bind_stream(), np_, ctypes.c_void_p(k.data_ptr()),
                         ctypes.c_void_p(kept.data_ptr())), "pic_select")
    ctx.synchronize()
    n = int(kept[0].item())
    return kept[1:1 + n]


def filter_particles(ctx, keep, *fields):
    """Keep the particles a creation functor accepted (keep[i] != 0), in
    order: filterEmpties (Picasso_ParticleInit.hpp:38-91) by pic_select +
    pic_permute.  Returns new [n_kept, q] tensors, one per field."""
    import torch
    idx = select(ctx, keep)
    outs = [torch.empty((f.shape[1], idx.shape[0]), dtype=f.dtype, device=f.device).t()
            for f in fields]
    if idx.shape[0] > 0:
        permute(ctx, idx, list(zip(fields, outs)))
    return outs


def bin_count(geom):
    return check(lib.pic_bin_count(ctypes.byref(geom.c())), "pic_bin_count")


def bin_sort(ctx, mesh, x, geom, out=None, with_active=True):
    """Stable counting sort of particles by bin key (brick-major)."""
    import torch
    nb = bin_count(geom)
    np_ = x.shape[0]
    if out is None:
        i32 = dict(dtype=torch.int32, device=x.device)
        out = BinSort(torch.empty(np_, **i32), torch.empty(np_, **i32),
                      torch.empty(nb + 1, **i32), geom,
                      torch.empty(nb + 1, **i32) if with_active else None)
    _check_bins(out, np_, nb)
    args = (ctypes.byref(mesh.c()), np_, ctypes.byref(_pfield(x, 3, "x")),
            ctypes.byref(geom.c()), out.keys.data_ptr(), out.perm.data_ptr(),
            out.bin_start.data_ptr(), out.active.data_ptr() if out.active is not None else None)
    check(lib.pic_bin_sort(ctx._bind_stream(), _dtype_code(x), *args), "pic_bin_sort")
    return out


def _check_bins(bins, np_, nb=None):
    """A BinSort must hold np_ keys and slots and nb + 1 brick offsets (and
    active entries); raises ValueError before anything is launched."""
    import torch
    if nb is None:
        nb = bin_count(bins.geom)
    parts = [(name, t, need) for name, t, need in
             (("perm", bins.perm, np_), ("keys", bins.keys, np_),
              ("bin_start", bins.bin_start, nb + 1), ("active", bins.active, nb + 1))
             if t is not None or name != "active"]
    for name, t, need in parts:
        if t is None:
            raise ValueError(f"bins.{name} is missing")
        if t.dtype not in (torch.int32, torch.uint32) or t.dim() != 1 or not t.is_contiguous():
            raise TypeError(f"bins.{name} must be a contiguous 1-D int32 tensor")
        if t.shape[0] < need:
            raise ValueError(f"bins.{name} holds {t.shape[0]} entries, fewer than {need}")
    for name, t, _ in parts:
        _require_device(t, f"bins.{name}")
