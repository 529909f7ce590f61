"""Host-side logic of the Python mirror, without a GPU: field validation
order (every shape / dtype check fires before anything touches a device)
and the AoSoA slice descriptor against the pic_pfield formula of
include/picasso_amd.h."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


def test_pfield_checks_precede_the_device_check(pa):
    from picasso_amd import _pfield
    t = torch.zeros((10, 3), dtype=torch.float64)
    with pytest.raises(ValueError, match=r"\[np, 12\]"):
        _pfield(t, 12, "c_p")
    with pytest.raises(ValueError, match="fewer than np"):
        _pfield(t, 3, "x", np_=11)
    with pytest.raises(TypeError, match="share x's dtype"):
        _pfield(t.float(), 3, "x", np_=10, dtype=pa.F64)
    with pytest.raises(ValueError, match="GPU tensor"):
        _pfield(t, 3, "x", np_=10, dtype=pa.F64)


def test_grid_checks(pa):
    from picasso_amd import _grid
    g = pa.GridField(torch.zeros((4, 4, 4, 3), dtype=torch.float32))
    with pytest.raises(TypeError):
        _grid(g, "mu_i", pa.F64, 3)
    with pytest.raises(ValueError, match="3 component"):
        _grid(pa.GridField(torch.zeros((4, 4, 4))), "mu_i", pa.F32, 3)
    with pytest.raises(ValueError, match="GPU tensor"):
        _grid(g, "mu_i", pa.F32, 3)


def test_bins_checks(pa):
    from picasso_amd import _check_bins
    geom = pa.Binning((0, 0, 0), (8, 8, 8), (4, 4, 4), 0)
    nb = pa.bin_count(geom)
    assert nb == 8
    i32 = dict(dtype=torch.int32)
    ok = pa.BinSort(torch.zeros(5, **i32), torch.zeros(5, **i32), torch.zeros(nb + 1, **i32), geom,
                    torch.zeros(nb + 1, **i32))
    with pytest.raises(ValueError, match="perm holds 5"):
        _check_bins(ok, 6)
    short = pa.BinSort(ok.perm, ok.keys, torch.zeros(nb, **i32), geom, None)
    with pytest.raises(ValueError, match="bin_start"):
        _check_bins(short, 5)
    wrong = pa.BinSort(ok.perm.long(), ok.keys, ok.bin_start, geom, None)
    with pytest.raises(TypeError):
        _check_bins(wrong, 5)
    with pytest.raises(ValueError, match="GPU tensor"):
        _check_bins(ok, 5)


@pytest.mark.parametrize("V", [1, 4, 16])
def test_aosoa_slice_formula(pa, V):
    """Element (p, c) of every member read back through the descriptor's
    (data offset, vlen_log2, tile_stride, cstride) equals the packed field,
    for a ragged last tile too."""
    rng = np.random.default_rng(V)
    P = 37
    fields = [torch.as_tensor(rng.normal(size=(P, q))) for q in (3, 1, 12)]
    aos = pa.AoSoA.from_fields(fields, V)
    buf = aos.buffer.numpy()
    esz = aos.buffer.element_size()
    base = aos.buffer.data_ptr()
    for k, f in enumerate(fields):
        s = aos.slice(k)
        assert torch.equal(s.to_tensor(), f)
        pf = s.c()
        off = (pf.data - base) // esz
        assert (1 << pf.vlen_log2) == V and pf.tile_stride == V * aos.q and pf.cstride == V
        got = np.empty((P, f.shape[1]))
        for p in range(P):
            for c in range(f.shape[1]):
                got[p, c] = buf[off + (p >> pf.vlen_log2) * pf.tile_stride + (p & (V - 1))
                                + c * pf.cstride]
        np.testing.assert_array_equal(got, f.numpy())


def test_aosoa_rejects_bad_shapes(pa):
    with pytest.raises(ValueError):
        pa.AoSoA(torch.zeros(10), 3, (1,), 2)  # vector length not a power of two
    with pytest.raises(ValueError):
        pa.AoSoA(torch.zeros(10), 4, (3,), 5)  # 2 tiles x 4 x 3 > 10 scalars
