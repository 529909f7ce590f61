"""pic_bin_sort on input that is already in bin order, nearly so, or not:
the identity fast path must give exactly the stable counting sort's result
(perm = identity, keys in place), and any descent -- inside a key-pass
workgroup of 1024 particles or across two of them -- must send the call
down the full path, bit-exact against the oracle."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import orpic  # noqa: E402

pytestmark = pytest.mark.gpu

KEY_WG = 1024  # particles per workgroup of the key pass


@pytest.fixture(scope="module")
def pa():
    import picasso_amd
    return picasso_amd


@pytest.fixture(scope="module")
def wl0():
    from picasso_amd import synth
    return synth.apic_uniform(32, 8)


def _geom(pa, wl, brick, dual):
    flags = pa.dual_dims(dual)
    cell_lo, ncell = wl.cell_region
    return pa.Binning(tuple(int(v) for v in cell_lo),
                      tuple(int(v) + f for v, f in zip(ncell, flags)), brick, dual)


def _oracle(pa, wl, x, geom):
    keys, _ = orpic.cell_keys(x, wl.low, wl.h, geom.cell_lo, geom.ncell, geom.brick, geom.dual)
    bvol = int(np.prod(geom.brick))
    perm, kstart = orpic.bin_sort(keys, pa.bin_count(geom) * bvol)
    return keys, perm, kstart[::bvol]


def _gpu(pa, wl, x, geom, out=None):
    import gpu_util as gu
    res = pa.bin_sort(gu.ctx(), pa.Mesh(wl.low, wl.h), gu.soa(x), geom, out=out)
    gu.ctx().synchronize()
    return (res, gu.host(res.perm).view(np.uint32).copy(),
            gu.host(res.keys).view(np.uint32).copy(),
            gu.host(res.bin_start).view(np.uint32).copy())


def _assert_matches(pa, wl, x, geom, out=None):
    keys, perm, bstart = _oracle(pa, wl, x, geom)
    res, gperm, gkeys, gstart = _gpu(pa, wl, x, geom, out)
    np.testing.assert_array_equal(gperm, perm)
    np.testing.assert_array_equal(gkeys, keys[perm])
    np.testing.assert_array_equal(gstart, bstart)
    import gpu_util as gu
    act = gu.host(res.active).view(np.uint32)
    nonempty = np.flatnonzero(np.diff(bstart.astype(np.int64)) > 0)
    assert act[0] == len(nonempty)
    np.testing.assert_array_equal(act[1:1 + len(nonempty)], nonempty)
    return res


@pytest.mark.parametrize("brick", [(4, 4, 4), (8, 8, 8), (2, 4, 8)])
@pytest.mark.parametrize("dual", [0, 1])
def test_presorted_input_gives_the_identity(pa, wl0, brick, dual):
    geom = _geom(pa, wl0, brick, dual)
    _, perm, _ = _oracle(pa, wl0, wl0.x, geom)
    xs = wl0.x[perm]
    _, gperm, _, _ = _gpu(pa, wl0, xs, geom)
    np.testing.assert_array_equal(gperm, np.arange(len(xs), dtype=np.uint32))
    _assert_matches(pa, wl0, xs, geom)


@pytest.mark.parametrize("where", ["inside", "boundary", "last"])
def test_one_descent_takes_the_full_path(pa, wl0, where):
    """Bin-ordered input with one adjacent pair swapped: inside a key-pass
    workgroup, across two of them (positions 1023 / 1024 of a workgroup
    pair), or at the very end."""
    geom = _geom(pa, wl0, (4, 4, 4), 1)
    keys, perm, _ = _oracle(pa, wl0, wl0.x, geom)
    xs = wl0.x[perm].copy()
    ks = keys[perm]
    P = len(xs)
    if where == "inside":
        cand = [q for q in range(5 * KEY_WG + 100, 6 * KEY_WG - 100) if ks[q] < ks[q + 1]]
    elif where == "boundary":
        cand = [q * KEY_WG - 1 for q in range(7, 20) if ks[q * KEY_WG - 1] < ks[q * KEY_WG]]
    else:
        cand = [q for q in range(P - 2, 0, -1) if ks[q] < ks[q + 1]]
    assert cand
    q = cand[0]
    xs[[q, q + 1]] = xs[[q + 1, q]]
    res = _assert_matches(pa, wl0, xs, geom)
    import gpu_util as gu
    assert not np.array_equal(gu.host(res.perm).view(np.uint32), np.arange(P, dtype=np.uint32))


def test_presorted_with_outsiders_at_the_tail(pa, wl0):
    """Outsiders carry the largest key, so bin-ordered input followed by
    particles outside the bins is still sorted: identity perm, the
    outsiders' keys 0xffffffff, and the call reports them."""
    import gpu_util as gu
    geom = pa.Binning((3, 2, 2), (31, 32, 32), (4, 4, 4), 0)  # cells at i = 2 left out
    cells = np.floor((wl0.x - wl0.low) / wl0.h).astype(np.int64)
    outside = cells[:, 0] < 3
    inside_x = wl0.x[~outside]
    _, perm, _ = _oracle(pa, wl0, inside_x, geom)
    xs = np.concatenate([inside_x[perm], wl0.x[outside]])
    res = pa.bin_sort(gu.ctx(), pa.Mesh(wl0.low, wl0.h), gu.soa(xs), geom)
    with pytest.raises(pa.PicError):
        gu.ctx().synchronize()
    gperm = gu.host(res.perm).view(np.uint32)
    gkeys = gu.host(res.keys).view(np.uint32)
    nin = len(inside_x)
    np.testing.assert_array_equal(gperm, np.arange(len(xs), dtype=np.uint32))
    assert (gkeys[nin:] == 0xFFFFFFFF).all() and (gkeys[:nin] < 0xFFFFFFFF).all()
    assert gu.host(res.bin_start).view(np.uint32)[-1] == nin


def test_alternating_calls_reset_the_flag(pa, wl0):
    """Sorted, shuffled, sorted again through one output set: each call
    decides its own path."""
    geom = _geom(pa, wl0, (4, 4, 4), 1)
    _, perm, _ = _oracle(pa, wl0, wl0.x, geom)
    xs = wl0.x[perm]
    out = _assert_matches(pa, wl0, xs, geom)
    _assert_matches(pa, wl0, wl0.x, geom, out)
    _assert_matches(pa, wl0, xs, geom, out)
