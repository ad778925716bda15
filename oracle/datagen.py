"""Synthetic-data protocol of the reference, restated on the CPU.

Test infrastructure only (see ``oracle/__init__.py``).

Pinning:
* ``GlibcRand`` is checked draw-for-draw against ``libc.rand`` through
  ctypes (tests/test_oracle_datagen.py) and against the published first
  draws of ``srand(0)``/``srand(1)`` (1804289383, 846930886, ...).
* ``dlarnv`` is checked bit-for-bit against LAPACK ``dlarnv`` exported by
  the OpenBLAS that ships with scipy (the library routine the reference
  reaches through ``LAPACKE_dlarnv``).
* ``zsort_locations`` / ``morton_order`` are checked against a literal
  transcription-free brute force (sorting explicit interleaved keys) and by
  invariants (the output is a permutation of the quantised input, sorted by
  Morton key).
* ``locations_obs_zsort_inplace`` is checked against a direct recursive
  partition on small inputs with forced ties.
"""
from __future__ import annotations

import math

import numpy as np

RAND_MAX = 2147483647
UINT16_MAX = 65535
UINT32_MAX = 4294967295


class GlibcRand:
    """glibc ``srand``/``rand`` (TYPE_3 additive-feedback generator).

    glibc's ``random_r`` with the default 128-byte state: a 31-word lagged
    Fibonacci generator r[i] = r[i-3] + r[i-31] (mod 2^32), seeded by the
    Park-Miller minimal standard LCG (16807, 2^31-1, Schrage's method),
    with the first 310 outputs discarded; ``rand()`` returns r >> 1.
    ``srand(0)`` behaves like ``srand(1)``.  The reference calls ``srand``
    and ``rand`` in GenerateXYLoc/uniform_distribution/shuffle
    (misc/compute/MLE_misc.c:753-792, 1238-1259).
    """

    def __init__(self, seed: int = 1):
        self.srand(seed)

    def srand(self, seed: int) -> None:
        seed &= 0xFFFFFFFF
        if seed == 0:
            seed = 1
        r = [0] * 34
        # The seed is stored as a signed 32-bit word.
        r[0] = seed - (1 << 32) if seed >= (1 << 31) else seed
        for i in range(1, 31):
            hi, lo = divmod(r[i - 1], 127773) if r[i - 1] >= 0 else (
                -((-r[i - 1]) // 127773), -((-r[i - 1]) % 127773))
            word = 16807 * lo - 2836 * hi
            if word < 0:
                word += 2147483647
            r[i] = word
        r = [v & 0xFFFFFFFF for v in r[:31]]
        # Ring of the last 31 words.  The first generated word is
        # r[34] = r[3] + r[31], with r[31..33] = r[0..2]: slot 3 holds
        # r[i-31] and slot (3+28) mod 31 = 0 holds r[i-3].
        self._ring = r
        self._pos = 3
        for _ in range(310):
            self._next_word()

    def _next_word(self) -> int:
        ring = self._ring
        p = self._pos
        # ring[p] is r[i-31]; r[i-3] is 28 positions ahead (mod 31).
        v = (ring[p] + ring[(p + 28) % 31]) & 0xFFFFFFFF
        ring[p] = v
        self._pos = (p + 1) % 31
        return v

    def rand(self) -> int:
        return self._next_word() >> 1


def uniform_distribution(rng: GlibcRand, lo: float, hi: float) -> float:
    """misc/compute/MLE_misc.c:753-761."""
    my_rand = float(rng.rand()) / (1.0 + RAND_MAX)
    return my_rand * (hi - lo) + lo


def _part1by1(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & 0xFFFF
    v = (v | (v << 8)) & 0x00FF00FF
    v = (v | (v << 4)) & 0x0F0F0F0F
    v = (v | (v << 2)) & 0x33333333
    v = (v | (v << 1)) & 0x55555555
    return v


def _compact1by1(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & 0x55555555
    v = (v | (v >> 1)) & 0x33333333
    v = (v | (v >> 2)) & 0x0F0F0F0F
    v = (v | (v >> 4)) & 0x00FF00FF
    v = (v | (v >> 8)) & 0x0000FFFF
    return v


def zsort_locations(x: np.ndarray, y: np.ndarray):
    """Morton sort with 16-bit quantisation (misc/compute/MLE_misc.c:297-319).

    Each coordinate is quantised as (uint16)(c*65535 + 0.5); the key puts
    the y bits in the odd positions; keys are sorted ascending and decoded
    back to c = q/65535.  Returns new (x, y) arrays.
    """
    qx = (np.asarray(x, dtype=np.float64) * float(UINT16_MAX) + 0.5).astype(np.uint64)
    qy = (np.asarray(y, dtype=np.float64) * float(UINT16_MAX) + 0.5).astype(np.uint64)
    code = (_part1by1(qy) << 1) + _part1by1(qx)
    code = np.sort(code.astype(np.uint32), kind="stable").astype(np.uint64)
    dx = _compact1by1(code)
    dy = _compact1by1(code >> 1)
    return dx.astype(np.float64) / float(UINT16_MAX), dy.astype(np.float64) / float(UINT16_MAX)


def generate_xy_loc(n: int, seed: int, rng: GlibcRand | None = None):
    """GenerateXYLoc (misc/compute/MLE_misc.c:763-792).

    Jittered grid on the unit square: sqrtn = ceil(sqrt(n)); for the i-th
    grid row and j-th column (both 1-based) x = (i - 0.5 + U(-0.4,0.4))/sqrtn,
    then y likewise with j, x drawn before y; truncated at n points; then
    Morton-sorted by ``zsort_locations``.  Returns (x, y, rng) where ``rng``
    holds the generator state after the 2n draws (the reference continues
    the same ``rand()`` stream in ``shuffle``).
    """
    if rng is None:
        rng = GlibcRand(seed)
    else:
        rng.srand(seed)
    sqrtn = int(math.ceil(math.sqrt(n)))
    x = np.empty(n)
    y = np.empty(n)
    idx = 0
    for i in range(sqrtn):
        if idx >= n:
            break
        for j in range(sqrtn):
            if idx >= n:
                break
            x[idx] = ((i + 1) - 0.5 + uniform_distribution(rng, -0.4, 0.4)) / sqrtn
            y[idx] = ((j + 1) - 0.5 + uniform_distribution(rng, -0.4, 0.4)) / sqrtn
            idx += 1
    x, y = zsort_locations(x, y)
    return x, y, rng


def generate_xy_loc_st(n: int, t_slots: int, seed: int, rng: GlibcRand | None = None):
    """GenerateXYLoc_ST (misc/compute/MLE_misc.c:887-923): the jittered grid
    of generate_xy_loc (x drawn before y, truncated at n) without the Morton
    sort, repeated for every time slot: location i + j n is spatial point i
    at time z = j + 1 (j = 0 .. t_slots - 1).  Returns (x, y, z, rng)."""
    if rng is None:
        rng = GlibcRand(seed)
    else:
        rng.srand(seed)
    sqrtn = int(math.ceil(math.sqrt(n)))
    x = np.empty(n)
    y = np.empty(n)
    idx = 0
    for i in range(sqrtn):
        if idx >= n:
            break
        for j in range(sqrtn):
            if idx >= n:
                break
            x[idx] = ((i + 1) - 0.5 + uniform_distribution(rng, -0.4, 0.4)) / sqrtn
            y[idx] = ((j + 1) - 0.5 + uniform_distribution(rng, -0.4, 0.4)) / sqrtn
            idx += 1
    z = np.repeat(np.arange(1.0, t_slots + 1.0), n)
    return np.tile(x, t_slots), np.tile(y, t_slots), z, rng


_TWOPI = 6.28318530717958647692528676655900576839
_DLARUV_A = 33952834046453
_M48 = (1 << 48) - 1
_DLARUV_MULT = [pow(_DLARUV_A, i, 1 << 48) for i in range(1, 129)]


def dlarnv(idist: int, iseed, n: int):
    """LAPACK dlarnv/dlaruv restated (the reference calls
    ``LAPACKE_dlarnv(3, {seed,seed,seed,1}, N, Nrand)``,
    examples/synthetic_dmle_test.c:111,147).

    dlaruv: multiplicative congruential generator modulo 2^48 with
    multiplier a = 33952834046453 (Fishman 1990); a call producing m <= 128
    numbers returns u_i = (s*a^i mod 2^48)/2^48, i = 1..m, and advances the
    seed to s*a^m.  dlarnv with idist=3 works in chunks of 64 normals, each
    from 128 uniforms by Box-Muller: sqrt(-2 log u1) cos(2 pi u2).
    Returns (values, new_iseed).
    """
    s = (int(iseed[0]) << 36) | (int(iseed[1]) << 24) | (int(iseed[2]) << 12) | int(iseed[3])
    out = np.empty(n)
    iv = 0
    while iv < n:
        il = min(64, n - iv)
        il2 = 2 * il if idist == 3 else il
        u = [((s * _DLARUV_MULT[i]) & _M48) / 281474976710656.0 for i in range(il2)]
        s = (s * _DLARUV_MULT[il2 - 1]) & _M48
        if idist == 1:
            out[iv:iv + il] = u
        elif idist == 2:
            out[iv:iv + il] = [2.0 * v - 1.0 for v in u]
        elif idist == 3:
            for i in range(il):
                out[iv + i] = math.sqrt(-2.0 * math.log(u[2 * i])) * math.cos(_TWOPI * u[2 * i + 1])
        else:
            raise ValueError("idist must be 1, 2 or 3")
        iv += il
    new_seed = [(s >> 36) & 4095, (s >> 24) & 4095, (s >> 12) & 4095, s & 4095]
    return out, new_seed


def shuffle(array: np.ndarray, x: np.ndarray, y: np.ndarray, rng: GlibcRand) -> None:
    """In-place Fisher-Yates variant of misc/compute/MLE_misc.c:1238-1259:
    j = i + rand() / (RAND_MAX/(n-i) + 1) with integer division; swaps the
    value and both coordinates."""
    n = len(array)
    if n <= 1:
        return
    for i in range(n - 1):
        j = i + rng.rand() // (RAND_MAX // (n - i) + 1)
        array[i], array[j] = array[j], array[i]
        x[i], x[j] = x[j], x[i]
        y[i], y[j] = y[j], y[i]


def _obs_keys(x: np.ndarray, y: np.ndarray):
    """Rescaled uint32 coordinates used by locations_obs_zsort_inplace
    (misc/compute/MLE_misc.c:476-500): (c - min)/(max - min) * UINT32_MAX
    truncated."""
    out = []
    for c in (x, y):
        c = np.asarray(c, dtype=np.float64)
        mn, mx = c.min(), c.max()
        out.append(((c - mn) / (mx - mn) * float(UINT32_MAX)).astype(np.uint32))
    return out


def obs_zsort_order(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Permutation produced by the hierarchical radix sort of
    locations_obs_zsort_inplace (misc/compute/MLE_misc.c:405-451,453-535).

    The partition visits bits from 31 down to 0 and, within a bit, the y
    coordinate before x; zero-bit elements keep their order at the low end,
    one-bit elements are written from the high end (order reversed).  The
    result is ascending order of the 64-bit interleaved key; inside a group
    of identical keys the input order is reversed once per set bit of the
    key, i.e. kept when popcount(key) is even and reversed when odd.
    """
    ux, uy = _obs_keys(x, y)
    n = len(ux)
    key = np.zeros(n, dtype=np.uint64)
    for b in range(31, -1, -1):
        key = (key << np.uint64(1)) | ((uy.astype(np.uint64) >> np.uint64(b)) & np.uint64(1))
        key = (key << np.uint64(1)) | ((ux.astype(np.uint64) >> np.uint64(b)) & np.uint64(1))
    pop = np.zeros(n, dtype=np.int64)
    kk = key.copy()
    while np.any(kk):
        pop += (kk & np.uint64(1)).astype(np.int64)
        kk >>= np.uint64(1)
    pos = np.arange(n, dtype=np.int64)
    tie_rank = np.where(pop % 2 == 0, pos, -pos)
    return np.lexsort((tie_rank, key))


def obs_zsort_order_recursive(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Direct partition recursion used to pin ``obs_zsort_order`` on small
    inputs (same visiting order: dims y then x for each bit 31..0)."""
    ux, uy = _obs_keys(x, y)
    data = [ux, uy]
    order = list(range(len(ux)))

    def rec(sdim, sbit, lo, hi):
        seg = order[lo:hi + 1]
        bit = 1 << sbit
        zeros = [e for e in seg if (int(data[sdim][e]) & bit) == 0]
        ones = [e for e in seg if (int(data[sdim][e]) & bit) != 0]
        order[lo:hi + 1] = zeros + ones[::-1]
        lo_last = lo + len(zeros)
        hi_last = hi - len(ones)
        if sdim > 0:
            nd, nb = sdim - 1, sbit
        elif sbit > 0:
            nd, nb = 1, sbit - 1
        else:
            return
        if lo_last - lo > 1:
            rec(nd, nb, lo, lo_last - 1)
        if hi - hi_last > 1:
            rec(nd, nb, hi_last + 1, hi)

    if len(order) > 1:
        rec(1, 31, 0, len(order) - 1)
    return np.array(order, dtype=np.int64)


def locations_obs_zsort_inplace(x: np.ndarray, y: np.ndarray, z: np.ndarray) -> None:
    """Reorder (x, y, z) in place by ``obs_zsort_order``."""
    order = obs_zsort_order(x, y)
    x[:] = x[order]
    y[:] = y[order]
    z[:] = z[order]


def pick_random_points(x: np.ndarray, y: np.ndarray, z: np.ndarray, n_miss: int,
                       rng: GlibcRand):
    """misc/compute/MLE_misc.c:1700-1741: shuffle (value, location) pairs
    with the continuing ``rand()`` stream, the first ``n_miss`` become the
    held-out set, the rest the observed set; both are then Morton-sorted
    with ``locations_obs_zsort_inplace``.  Returns
    (xobs, yobs, zobs, xmiss, ymiss, zactual)."""
    zz = np.array(z, dtype=np.float64, copy=True)
    xx = np.array(x, dtype=np.float64, copy=True)
    yy = np.array(y, dtype=np.float64, copy=True)
    shuffle(zz, xx, yy, rng)
    xm, ym, zm = xx[:n_miss].copy(), yy[:n_miss].copy(), zz[:n_miss].copy()
    xo, yo, zo = xx[n_miss:].copy(), yy[n_miss:].copy(), zz[n_miss:].copy()
    locations_obs_zsort_inplace(xo, yo, zo)
    locations_obs_zsort_inplace(xm, ym, zm)
    return xo, yo, zo, xm, ym, zm
