"""pic_bin_sort_noalloc through the C ABI against pic_bin_sort: bit-exact
keys, permutation, brick offsets and active list on two calls in a row (the
second smaller, so stale contents of the shared device array would show),
and a clear PIC_ERR_CAPACITY beyond the array's size."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import orpic  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


def _call(pa, fn, ctx, wl, x, geom, np_):
    """One sort through the C ABI entry point `fn` into fresh buffers."""
    from picasso_amd import _abi
    dev = "cuda:0"
    nb = pa.bin_count(geom)
    keys = torch.full((np_,), -7, dtype=torch.int32, device=dev)
    perm = torch.full((np_,), -7, dtype=torch.int32, device=dev)
    start = torch.full((nb + 1,), -7, dtype=torch.int32, device=dev)
    act = torch.full((nb + 1,), -7, dtype=torch.int32, device=dev)
    xf = pa._pfield(x, 3, "x")
    rc = fn(ctx._bind_stream(), _abi.F64, ctypes.byref(pa.Mesh(wl.low, wl.h).c()), np_,
            ctypes.byref(xf), ctypes.byref(geom.c()), keys.data_ptr(), perm.data_ptr(),
            start.data_ptr(), act.data_ptr())
    pa.check(rc, fn.__name__)
    ctx.synchronize()
    return [t.cpu().numpy().view(np.uint32) for t in (keys, perm, start, act)]


def test_noalloc_sort_matches_the_allocating_sort(pa):
    from picasso_amd import _abi, synth
    import gpu_util as gu
    wl = synth.apic_uniform(32, 8)
    ctx = gu.ctx()
    cell_lo, ncell = wl.cell_region
    geom = pa.Binning(tuple(int(v) for v in cell_lo), tuple(int(v) + 1 for v in ncell),
                      (4, 4, 4), 1)
    keys_ref, _ = orpic.cell_keys(wl.x, wl.low, wl.h, geom.cell_lo, geom.ncell, (4, 4, 4), 1)
    # first call: the whole shuffled workload; second: a smaller, bin-ordered
    # prefix of it (the identity path) with a different binning
    order = np.argsort(keys_ref, kind="stable")
    cases = [(wl.x, geom),
             (wl.x[order][:50000],
              pa.Binning(tuple(int(v) for v in cell_lo), tuple(int(v) + 1 for v in ncell),
                         (8, 8, 8), 1))]
    for xh, g in cases:
        x = gu.soa(xh)
        got = _call(pa, _abi.lib.pic_bin_sort_noalloc, ctx, wl, x, g, len(xh))
        ref = _call(pa, _abi.lib.pic_bin_sort, ctx, wl, x, g, len(xh))
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
        nact = int(ref[3][0])
        assert nact > 0 and int(ref[2][-1]) == len(xh)


def test_noalloc_sort_rejects_oversize(pa):
    """4,194,304 particles fit the array with any binning below 2^20
    bricks; 6M particles do not: PIC_ERR_CAPACITY, nothing launched."""
    from picasso_amd import _abi
    import gpu_util as gu
    P = 6_000_000
    x = torch.zeros((3, P), dtype=torch.float64, device="cuda:0").t()
    geom = pa.Binning((0, 0, 0), (8, 8, 8), (4, 4, 4), 0)
    nb = pa.bin_count(geom)
    i32 = dict(dtype=torch.int32, device="cuda:0")
    keys, perm = torch.empty(P, **i32), torch.empty(P, **i32)
    start, act = torch.empty(nb + 1, **i32), torch.empty(nb + 1, **i32)
    rc = _abi.lib.pic_bin_sort_noalloc(gu.ctx()._bind_stream(), _abi.F64,
                                       ctypes.byref(pa.Mesh((0, 0, 0), 0.125).c()), P,
                                       ctypes.byref(pa._pfield(x, 3, "x")),
                                       ctypes.byref(geom.c()), keys.data_ptr(), perm.data_ptr(),
                                       start.data_ptr(), act.data_ptr())
    assert rc == _abi.PIC_ERR_CAPACITY
    gu.ctx().synchronize()
