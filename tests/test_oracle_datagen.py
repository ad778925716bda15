"""Pins of the oracle's synthetic-data protocol and the library's host-side
generators (CPU only)."""
import ctypes
import glob
import os

import numpy as np
import pytest

from oracle import datagen

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _libc():
    return ctypes.CDLL("libc.so.6")


def _openblas():
    import scipy
    libs = glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs", "libscipy_openblas*.so"))
    if not libs:
        pytest.skip("scipy's bundled OpenBLAS not found")
    return ctypes.CDLL(libs[0])


@pytest.mark.parametrize("seed", [0, 1, 2, 12345, 2**31 + 5, 2**32 - 1])
def test_glibc_rand_matches_libc(seed):
    libc = _libc()
    libc.srand(ctypes.c_uint(seed))
    g = datagen.GlibcRand(seed)
    assert [libc.rand() for _ in range(2000)] == [g.rand() for _ in range(2000)]


def test_glibc_rand_published_values():
    g = datagen.GlibcRand(1)
    assert [g.rand() for _ in range(3)] == [1804289383, 846930886, 1681692777]
    assert datagen.GlibcRand(0).rand() == 1804289383  # srand(0) == srand(1)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 1601])
@pytest.mark.parametrize("seed", [0, 1, 9])
def test_dlarnv_matches_lapack(n, seed):
    lib = _openblas()
    iseed = (ctypes.c_int * 4)(seed, seed, seed, 1)
    ref = np.zeros(n)
    lib.scipy_dlarnv_(ctypes.byref(ctypes.c_int(3)), iseed, ctypes.byref(ctypes.c_int(n)),
                      ref.ctypes.data_as(ctypes.c_void_p))
    val, s = datagen.dlarnv(3, [seed, seed, seed, 1], n)
    assert np.array_equal(val, ref)
    assert s == list(iseed)


def test_dlarnv_uniform_modes():
    lib = _openblas()
    for idist in (1, 2):
        iseed = (ctypes.c_int * 4)(3, 4, 5, 7)
        ref = np.zeros(300)
        lib.scipy_dlarnv_(ctypes.byref(ctypes.c_int(idist)), iseed, ctypes.byref(ctypes.c_int(300)),
                          ref.ctypes.data_as(ctypes.c_void_p))
        val, _ = datagen.dlarnv(idist, [3, 4, 5, 7], 300)
        assert np.array_equal(val, ref)


def test_zsort_is_sorted_permutation_of_quantised_input():
    rng = np.random.default_rng(1)
    x = rng.random(500)
    y = rng.random(500)
    xs, ys = datagen.zsort_locations(x, y)
    qx = np.floor(x * 65535 + 0.5)
    qy = np.floor(y * 65535 + 0.5)
    # same multiset of quantised points
    a = sorted(zip(qx.tolist(), qy.tolist()))
    b = sorted(zip((xs * 65535).round().tolist(), (ys * 65535).round().tolist()))
    assert a == b

    # keys ascend; brute-force interleave (y bit above x bit)
    def key(px, py):
        k = 0
        for bit in range(15, -1, -1):
            k = (k << 1) | ((int(py) >> bit) & 1)
            k = (k << 1) | ((int(px) >> bit) & 1)
        return k
    keys = [key(round(u * 65535), round(v * 65535)) for u, v in zip(xs, ys)]
    assert keys == sorted(keys)


def test_generate_xy_loc_golden():
    g = np.load(os.path.join(GOLD, "datagen_small.npz"))
    for n, seed in [(10, 0), (100, 1), (1000, 3), (1600, 0)]:
        x, y, _ = datagen.generate_xy_loc(n, seed)
        assert np.array_equal(x, g[f"x_{n}_{seed}"]) and np.array_equal(y, g[f"y_{n}_{seed}"])
    # on the 40 x 40 grid every point stays inside its jitter cell
    x, y, _ = datagen.generate_xy_loc(1600, 0)
    cx = np.floor(x * 40)
    cy = np.floor(y * 40)
    assert len(set(zip(cx.tolist(), cy.tolist()))) == 1600


def test_generate_xy_loc_truncated_grid():
    x, y, _ = datagen.generate_xy_loc(50, 0)  # ceil(sqrt(50)) = 8: 6 full rows + 2
    assert len(x) == 50 and np.all((x > 0) & (x < 1)) and np.all((y > 0) & (y < 1))


@pytest.mark.parametrize("seed", [0, 4])
def test_obs_zsort_matches_partition_sort(seed):
    rng = np.random.default_rng(seed)
    n = 300
    x = rng.random(n)
    y = rng.random(n)
    # force exact ties (identical points) and shared coordinates
    x[10:15] = x[3]
    y[10:15] = y[3]
    x[40:50] = 0.5
    order = datagen.obs_zsort_order(x, y)
    ref = datagen.obs_zsort_order_recursive(x, y)
    assert np.array_equal(order, ref)


def test_shuffle_uses_continuing_stream():
    rng = datagen.GlibcRand(0)
    a = np.arange(10.0)
    x = np.arange(10.0)
    y = np.arange(10.0)
    datagen.shuffle(a, x, y, rng)
    assert sorted(a.tolist()) == list(range(10))
    assert np.array_equal(a, x) and np.array_equal(a, y)
    libc = _libc()
    libc.srand(0)
    b = list(range(10))
    for i in range(9):
        j = i + libc.rand() // (2147483647 // (10 - i) + 1)
        b[i], b[j] = b[j], b[i]
    assert a.tolist() == [float(v) for v in b]


def test_library_generate_xy_loc_matches_oracle():
    """The C ABI's GenerateXYLoc (libc rand + Morton sort) equals the oracle
    bit for bit; host-only, no GPU call."""
    import exageostat_amd
    try:
        exageostat_amd.load()
    except RuntimeError as exc:
        pytest.skip(str(exc))
    for n, seed in [(10, 0), (1000, 3), (1600, 0), (2025, 7)]:
        x, y = exageostat_amd.generate_xy_loc(n, seed)
        xo, yo, _ = datagen.generate_xy_loc(n, seed)
        assert np.array_equal(x, xo) and np.array_equal(y, yo)


def test_generate_xy_loc_st_shares_the_grid_stream():
    """GenerateXYLoc_ST draws the same jittered grid as GenerateXYLoc (same
    rand() stream, x before y): Morton-sorting its first slot gives
    GenerateXYLoc's points bit for bit; every slot repeats them with time
    slot + 1."""
    for n, seed, slots in [(100, 1, 3), (1600, 0, 2), (50, 5, 4)]:
        x, y, z, _ = datagen.generate_xy_loc_st(n, slots, seed)
        xs, ys = datagen.zsort_locations(x[:n], y[:n])
        xo, yo, _ = datagen.generate_xy_loc(n, seed)
        assert np.array_equal(xs, xo) and np.array_equal(ys, yo)
        for t in range(slots):
            assert np.array_equal(x[t * n:(t + 1) * n], x[:n]) and np.all(z[t * n:(t + 1) * n] == t + 1)


def test_library_generate_xy_loc_st_matches_oracle():
    import exageostat_amd
    try:
        exageostat_amd.load()
    except RuntimeError as exc:
        pytest.skip(str(exc))
    for n, slots, seed in [(100, 3, 1), (1000, 2, 0)]:
        x, y, z = exageostat_amd.generate_xy_loc_st(n, slots, seed)
        xo, yo, zo, _ = datagen.generate_xy_loc_st(n, slots, seed)
        assert np.array_equal(x, xo) and np.array_equal(y, yo) and np.array_equal(z, zo)
