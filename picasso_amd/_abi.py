"""ctypes view of the C ABI declared in include/picasso_amd.h.

Loads picasso_amd/libpicasso_amd.so (built by picasso_amd/build.py).  There
is no fallback: if the library is missing or fails to load, importing this
module raises, so nothing can silently run a non-HIP path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpicasso_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "picasso_amd.h")

PIC_OK = 0
PIC_ERR_ARG = -1
PIC_ERR_BOUNDS = -2
PIC_ERR_HIP = -3
PIC_ERR_ALLOC = -4
PIC_ERR_UNSUPPORTED = -5
PIC_ERR_CAPACITY = -6

NODE, CELL, FACE_I, FACE_J, FACE_K, EDGE_I, EDGE_J, EDGE_K = range(8)
F64, F32 = 0, 1
P2G_ZERO_FIRST = 1
P2G_DETERMINISTIC = 2
OPT_G2P_GROUP = 1  # pic_set_option: work-items per collocated g2p brick workgroup


class Mesh(ctypes.Structure):
    _fields_ = [("low", ctypes.c_double * 3), ("h", ctypes.c_double)]


class Grid(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("lo", ctypes.c_int32 * 3),
                ("n", ctypes.c_int32 * 3), ("stride", ctypes.c_int64 * 3),
                ("cstride", ctypes.c_int64)]


class PField(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("vlen_log2", ctypes.c_int32),
                ("tile_stride", ctypes.c_int64), ("cstride", ctypes.c_int64)]


class Binning(ctypes.Structure):
    _fields_ = [("cell_lo", ctypes.c_int32 * 3), ("ncell", ctypes.c_int32 * 3),
                ("brick", ctypes.c_int32 * 3), ("dual", ctypes.c_int32)]


class Bins(ctypes.Structure):
    _fields_ = [("perm", ctypes.c_void_p), ("keys", ctypes.c_void_p),
                ("bin_start", ctypes.c_void_p), ("active", ctypes.c_void_p),
                ("geom", Binning)]


class Instance(ctypes.Structure):
    _fields_ = [("mesh", Mesh), ("first", ctypes.c_int64), ("np", ctypes.c_int64),
                ("geom", Binning), ("m_i", Grid), ("mu_i", Grid)]


class FieldDep(ctypes.Structure):
    _fields_ = [("grid", Grid), ("ncomp", ctypes.c_int32), ("own_lo", ctypes.c_int32 * 3),
                ("own_n", ctypes.c_int32 * 3), ("periodic", ctypes.c_int32 * 3)]


APPLY_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)

_P = ctypes.POINTER
_c_int, _c_i64, _c_u32, _c_double, _vp = (ctypes.c_int, ctypes.c_int64, ctypes.c_uint32,
                                          ctypes.c_double, ctypes.c_void_p)
_I32x3 = ctypes.c_int32 * 3

# name -> (restype, argtypes); every function of include/picasso_amd.h.
PROTOTYPES = {
    "pic_context_create": (_c_int, [_P(_vp), _c_int]),
    "pic_context_destroy": (_c_int, [_vp]),
    "pic_set_stream": (_c_int, [_vp, _vp]),
    "pic_synchronize": (_c_int, [_vp]),
    "pic_status_string": (ctypes.c_char_p, [_c_int]),
    "pic_api_version": (_c_int, []),
    "pic_apic_p2g": (_c_int, [_vp, _c_int, _c_int, _c_int, _P(Mesh), _c_i64, _P(PField),
                              _P(PField), _P(PField), _P(Grid), _P(Grid), _P(Bins), _c_u32]),
    "pic_apic_g2p": (_c_int, [_vp, _c_int, _c_int, _c_int, _P(Mesh), _c_i64, _P(PField),
                              _P(Grid), _P(PField), _P(Bins)]),
    "pic_polypic_nmode": (_c_int, [_c_int]),
    "pic_polypic_p2g": (_c_int, [_vp, _c_int, _c_int, _c_int, _P(Mesh), _c_i64, _c_int,
                                 _P(PField), _P(PField), _P(PField), _P(PField), _P(Grid),
                                 _P(Grid), _c_double, _P(Bins), _c_u32]),
    "pic_polypic_g2p": (_c_int, [_vp, _c_int, _c_int, _c_int, _P(Mesh), _c_i64, _c_int,
                                 _P(PField), _P(Grid), _P(PField), _P(Bins)]),
    "pic_grid_normalize": (_c_int, [_vp, _c_int, _c_int, _P(Grid), _P(Grid), _P(Grid),
                                    _c_double]),
    "pic_g2p_gradu": (_c_int, [_vp, _c_int, _c_int, _c_int, _P(Mesh), _c_i64, _P(PField),
                               _P(Grid), _P(PField), _P(Bins)]),
    "pic_g2p_interp": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _P(Mesh), _c_i64,
                                _P(PField), _P(Grid), _c_int, _P(PField), _P(Bins)]),
    "pic_p2g_interp": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _P(Mesh), _c_i64,
                                _P(PField), _P(PField), _c_int, _P(Grid), _c_u32, _P(Bins)]),
    "pic_update_F": (_c_int, [_vp, _c_i64, _c_int, _P(PField), _c_double, _P(PField),
                              _P(PField), _P(PField)]),
    "pic_bin_count": (_c_i64, [_P(Binning)]),
    "pic_bin_sort": (_c_int, [_vp, _c_int, _P(Mesh), _c_i64, _P(PField), _P(Binning), _vp,
                              _vp, _vp, _vp]),
    "pic_bin_sort_noalloc": (_c_int, [_vp, _c_int, _P(Mesh), _c_i64, _P(PField), _P(Binning),
                                      _vp, _vp, _vp, _vp]),
    "pic_apic_step": (_c_int, [_vp, _c_int, _c_int, _c_int, _P(Mesh), _c_i64, _P(PField),
                               _P(PField), _P(PField), _P(Grid), _P(Grid), _P(Binning), _vp,
                               _vp, _vp, _vp]),
    "pic_apic_g2p_normalized": (_c_int, [_vp, _c_int, _c_int, _c_int, _P(Mesh), _c_i64,
                                         _P(PField), _P(Grid), _P(Grid), _P(PField), _P(Bins)]),
    "pic_apic_g2p_update_F": (_c_int, [_vp, _c_int, _c_int, _c_int, _P(Mesh), _c_i64,
                                       _P(PField), _P(Grid), _P(Grid), _P(PField), _c_double,
                                       _P(PField), _P(PField), _P(PField), _P(Bins)]),
    "pic_permute": (_c_int, [_vp, _c_int, _c_i64, _vp, _c_int, _P(PField), _P(PField),
                             _P(_c_int)]),
    "pic_init_particles": (_c_i64, [_vp, _c_int, _c_int, _P(Mesh), _P(ctypes.c_int32),
                                    _P(ctypes.c_int32), _c_int, ctypes.c_uint64, _c_i64,
                                    _P(PField), _P(PField)]),
    "pic_select": (_c_int, [_vp, _c_i64, _vp, _vp]),
    "pic_grid_halo": (_c_int, [_vp, _c_int, _c_int, _P(Grid), _c_int, _P(ctypes.c_int32),
                               _P(ctypes.c_int32), _P(ctypes.c_int32)]),
    "pic_batch_create": (_c_int, [_vp, _c_int, _c_int, _c_int, _P(Instance), _P(_vp)]),
    "pic_batch_destroy": (_c_int, [_vp]),
    "pic_batch_bin_sort": (_c_int, [_vp, _vp, _P(PField)]),
    "pic_batch_apic_p2g": (_c_int, [_vp, _vp, _P(PField), _P(PField), _P(PField), _c_u32]),
    "pic_batch_apic_g2p_normalized": (_c_int, [_vp, _vp, _P(PField), _P(PField)]),
    "pic_batch_apic_step": (_c_int, [_vp, _vp, _P(PField), _P(PField), _P(PField), _c_u32]),
    "pic_grid_operator_apply": (_c_int, [_vp, _c_int, _c_int, _P(FieldDep), _c_int,
                                         _P(FieldDep), APPLY_FN, _vp]),
    "pic_set_option": (_c_int, [_vp, _c_int, _c_int]),
    "pic_graph_begin": (_c_int, [_vp]),
    "pic_graph_end": (_c_int, [_vp, _P(_vp)]),
    "pic_graph_launch": (_c_int, [_vp, _vp]),
    "pic_graph_destroy": (_c_int, [_vp]),
    "pic_set_timing": (_c_int, [_vp, _c_int]),
    "pic_kernel_time": (_c_int, [_vp, ctypes.c_char_p, _P(_c_double), _P(_c_i64)]),
    "pic_reset_timing": (_c_int, [_vp]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -m picasso_amd.build or __graft_entry__.build())")
    # torch first: its HIP runtime (same soname) then serves this library
    # too.  Loaded the other way round, the library starts the system HIP
    # runtime and torch's own copy finds no GPU.
    import torch  # noqa: F401
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class PicError(RuntimeError):
    def __init__(self, code, what):
        self.code = code
        msg = lib.pic_status_string(code).decode()
        super().__init__(f"{what}: {msg} ({code})")


def check(rc, what):
    if rc < 0:
        raise PicError(rc, what)
    return rc
