"""The C-ABI library loads and exports every function include/picasso_amd.h
declares; host-only entry points behave (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "picasso_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(pic_[A-Za-z_0-9]+)\s*\(",
                                 text, flags=re.M)))


def test_header_declares_the_abi():
    names = declared_functions()
    for must in ("pic_apic_p2g", "pic_apic_g2p", "pic_polypic_p2g", "pic_polypic_g2p",
                 "pic_grid_normalize", "pic_bin_sort", "pic_update_F", "pic_g2p_gradu"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from picasso_amd import _abi
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes prototypes cover the same set
    assert sorted(_abi.PROTOTYPES) == declared_functions()


def test_host_only_entry_points():
    from picasso_amd import _abi
    lib = _abi.lib
    assert lib.pic_api_version() == 1
    assert lib.pic_status_string(-2) == b"stencil or cell outside its window"
    assert lib.pic_polypic_nmode(1) == 8
    assert lib.pic_polypic_nmode(3) == 27
    assert lib.pic_polypic_nmode(4) == _abi.PIC_ERR_ARG
    i3 = ctypes.c_int32 * 3

    def count(ncell, brick, dual=0):
        g = _abi.Binning(i3(0, 0, 0), i3(*ncell), i3(*brick), dual)
        return lib.pic_bin_count(ctypes.byref(g))

    assert count((256, 256, 256), (8, 8, 8)) == 32 ** 3
    assert count((257, 257, 257), (8, 8, 8), 1) == 33 ** 3
    assert count((10, 10, 10), (4, 4, 4)) == 27
    assert count((0, 10, 10), (4, 4, 4)) == _abi.PIC_ERR_ARG
    assert count((10, 10, 10), (4, 4, 4), 2) == _abi.PIC_ERR_ARG
    assert count((10, 10, 10), (4, 4, 4), 0x10 | 5) == 27
    assert count((10, 10, 10), (4, 4, 4), 0x10) == 27
    assert count((10, 10, 10), (4, 4, 4), 0x18) == _abi.PIC_ERR_ARG
    assert count((10, 10, 10), (4, 4, 4), 0x20 | 1) == _abi.PIC_ERR_ARG
    assert count((64, 64, 64), (32, 32, 8)) == _abi.PIC_ERR_CAPACITY


def test_null_context_is_an_argument_error():
    from picasso_amd import _abi
    lib = _abi.lib
    mesh = _abi.Mesh((ctypes.c_double * 3)(0, 0, 0), 0.1)
    assert lib.pic_apic_p2g(None, 2, 0, 0, ctypes.byref(mesh), 0, None, None, None, None, None,
                            None, 0) == _abi.PIC_ERR_ARG
    assert lib.pic_synchronize(None) == _abi.PIC_ERR_ARG
    assert lib.pic_context_destroy(None) == _abi.PIC_ERR_ARG


def test_no_fallback_when_library_missing(tmp_path, monkeypatch):
    """The package refuses to import without its HIP library."""
    import importlib.util
    src = os.path.join(ROOT, "picasso_amd", "_abi.py")
    spec = importlib.util.spec_from_file_location("abi_copy", src)
    mod = importlib.util.module_from_spec(spec)
    monkeypatch.setattr(os.path, "exists", lambda p: False if p.endswith(".so") else True)
    with pytest.raises(ImportError):
        spec.loader.exec_module(mod)


CPP_PROBE = r"""
#include "picasso_amd.hpp"
#include <cstdio>
int main() {
    namespace pa = picasso_amd;
    pic_binning g{{0, 0, 0}, {10, 10, 10}, {4, 4, 4}, 0};
    pic_mesh mesh{{0.0, 0.0, 0.0}, 0.5};
    double buf[12];
    pic_pfield x = pa::soa(buf, 4);
    pa::SplineData sd = pa::createSpline(PIC_NODE, 2, PIC_F64, mesh, 4, x);
    int lo[3] = {0, 0, 0}, n[3] = {4, 4, 4};
    pic_grid gm = pa::dense_grid(buf, lo, n, 1);
    try {
        pa::APIC::p2g(nullptr, x, x, gm, gm, sd);  // null context: argument error
        return 2;
    } catch (const pa::Error& e) {
        std::printf("%d %lld %d\n", e.status, (long long)pic_bin_count(&g), pic_api_version());
    }
    return 0;
}
"""


def test_cpp_facade_compiles_links_and_reports_errors(tmp_path):
    """The C++ facade builds with a plain host compiler against the library
    and turns a pic_status into an exception (no GPU call is made)."""
    import shutil
    import subprocess
    from picasso_amd import _abi
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "probe.cpp"
    src.write_text(CPP_PROBE)
    exe = tmp_path / "probe"
    libdir = os.path.dirname(_abi.LIB_PATH)
    subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe), "-L", libdir, "-lpicasso_amd",
                    "-Wl,-rpath," + libdir], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert out == [str(_abi.PIC_ERR_ARG), "27", "1"]
