"""EXAGEOSTAT_dtrsv_Tile_noalloc (the triangular solve that allocates no
device memory; its scratch is a static device array) against
EXAGEOSTAT_dtrsv_Tile on the same factor and right-hand side, through the
C ABI: two calls in a row, the second on a smaller matrix, forward and
transposed solves; the same kernels run in the same order, so the results
agree to the last bit (asserted at 1e-15 relative).  Sizes beyond the
array's limit are refused with an error."""
import ctypes

import numpy as np
import pytest

from oracle import datagen, matern

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    import exageostat_amd
    exageostat_amd.load()
    assert exageostat_amd.device_count() >= 1, "no HIP device visible"
    return exageostat_amd


def _solve_both(lib, n, ts, seed):
    from exageostat_amd import _lib
    c = _lib.load()
    x, y, _ = datagen.generate_xy_loc(n, seed)
    M = lib.DeviceMatrix(n, n, ts, lower=True)
    M.dcmg((1.0, 0.1, 0.5), x, y)
    assert M.potrf() == 0
    b = np.random.default_rng(seed).standard_normal(n)
    out = []
    for trans in (_lib.ChamNoTrans, _lib.ChamTrans):
        ref = b.copy()
        assert c.EXAGEOSTAT_dtrsv_Tile(trans, M.desc, ref.ctypes.data_as(_lib.c_double_p)) == 0
        got = b.copy()
        rc = c.EXAGEOSTAT_dtrsv_Tile_noalloc(trans, M.desc, got.ctypes.data_as(_lib.c_double_p))
        assert rc == 0, c.exageostat_gpu_last_error()
        out.append((ref, got))
    # the solves against the oracle's dense factor as well
    L = np.linalg.cholesky(matern.core_dcmg(x, y, x, y, (1.0, 0.1, 0.5)))
    assert np.allclose(L @ out[0][1], b, rtol=0, atol=1e-9 * np.max(np.abs(b)))
    M.close()
    return out


def test_noalloc_solve_matches_allocating_solve_twice(lib):
    print("noalloc: n=6000 ts=320, then n=1500 ts=128", flush=True)
    for n, ts, seed in [(6000, 320, 1), (1500, 128, 2)]:
        for ref, got in _solve_both(lib, n, ts, seed):
            assert np.max(np.abs(got - ref)) <= 1e-15 * np.max(np.abs(ref)), (n, ts)


def test_noalloc_solve_refuses_oversized_input(lib):
    from exageostat_amd import _lib
    c = _lib.load()
    n, ts = 65537, 1024  # mt * ts = 66,560 > 65,536 (18 GB of lower tiles, allocated, never computed on)
    M = lib.DeviceMatrix(n, n, ts, lower=True)
    b = np.zeros(n)
    rc = c.EXAGEOSTAT_dtrsv_Tile_noalloc(_lib.ChamNoTrans, M.desc, b.ctypes.data_as(_lib.c_double_p))
    assert rc == -1 and b"65536" in c.exageostat_gpu_last_error()
    M.close()
