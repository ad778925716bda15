"""Writes the golden fixtures under tests/golden/ (run from the repo root:
``python tests/golden/make_fixtures.py``).

The reference ships no golden vectors for this path, so the fixtures are
produced by the CPU oracle after it has been pinned against independent
checks (libc rand, LAPACK dlarnv, mpmath, scipy); the pins are exercised by
tests/test_oracle_*.py.  Files are plain .npz (no pickles) plus a JSON
manifest of SHA-256 digests.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import datagen, likelihood, matern  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
K = "univariate_matern_stationary"


def save(name: str, **arrays):
    path = os.path.join(OUT, name)
    np.savez(path, **arrays)
    return path


def knu_table():
    import mpmath as mp
    mp.mp.dps = 30
    nus = [0.1, 0.25, 0.5, 0.73, 1.0, 1.37, 1.5, 2.0, 2.5, 3.3, 4.9]
    xs = [1e-5, 1e-3, 0.05, 0.3, 1.0, 1.7, 2.0, 2.3, 4.0, 8.0, 16.0, 40.0]
    nu_col, x_col, k_col, k1_col, dk_col = [], [], [], [], []
    for nu in nus:
        for x in xs:
            nu_col.append(nu)
            x_col.append(x)
            k_col.append(float(mp.besselk(nu, x)))
            k1_col.append(float(mp.besselk(nu + 1, x)))
            dk_col.append(float(mp.diff(lambda v: mp.besselk(v, x), nu)))
    return save("knu_table.npz", nu=np.array(nu_col), x=np.array(x_col), k=np.array(k_col),
                k1=np.array(k1_col), dk_dnu=np.array(dk_col))


def datagen_small():
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(0)
    rand0 = np.array([libc.rand() for _ in range(20)], dtype=np.int64)
    out = {"rand_seed0": rand0}
    for n, seed in [(10, 0), (100, 1), (1000, 3), (1600, 0)]:
        x, y, _ = datagen.generate_xy_loc(n, seed)
        out[f"x_{n}_{seed}"] = x
        out[f"y_{n}_{seed}"] = y
    for n, seed in [(10, 0), (200, 1), (1600, 0)]:
        e, s = datagen.dlarnv(3, [seed, seed, seed, 1], n)
        out[f"e_{n}_{seed}"] = e
        out[f"eseed_{n}_{seed}"] = np.array(s, dtype=np.int64)
    return save("datagen_small.npz", **out)


def c1(seed: int = 0):
    """Config C1: n=1600 jittered grid, theta=(1, 0.1, 0.5), z = L e."""
    n = 1600
    x, y, rng = datagen.generate_xy_loc(n, seed)
    e, _ = datagen.dlarnv(3, [seed, seed, seed, 1], n)
    theta = np.array([1.0, 0.1, 0.5])
    z = likelihood.dzvg(K, x, y, theta, e)
    thetas = np.array([[1.0, 0.1, 0.5], [0.5, 0.05, 1.0], [2.0, 0.3, 0.8], [1.3, 0.07, 1.9], [0.8, 0.2, 2.5]])
    ll = np.array([likelihood.dmle(K, x, y, z, t) for t in thetas])
    # prediction split of the reference: shuffle with the continuing rand()
    xo, yo, zo, xm, ym, za = datagen.pick_random_points(x, y, z, 100, rng)
    zmiss, mse, _, _ = likelihood.predict(K, xo, yo, zo, xm, ym, theta, za)
    return save(f"c1_seed{seed}.npz", x=x, y=y, e=e, z=z, theta_true=theta, thetas=thetas,
                loglik=ll[:, 0], logdet=ll[:, 1], dotp=ll[:, 2], xobs=xo, yobs=yo, zobs=zo, xmiss=xm,
                ymiss=ym, zactual=za, zmiss=zmiss, mse=np.array([mse]))


def nugget_small():
    """Nugget kernel on 900 points: l and gradient at (1, 0.1, 0.5, 0.01)."""
    n = 900
    x, y, _ = datagen.generate_xy_loc(n, 5)
    e, _ = datagen.dlarnv(3, [5, 5, 5, 1], n)
    theta = np.array([1.0, 0.1, 0.5, 0.01])
    kern = "univariate_matern_nuggets_stationary"
    z = likelihood.dzvg(kern, x, y, theta, e)
    l, g = likelihood.gradient(kern, x, y, z, theta)
    return save("nugget_900.npz", x=x, y=y, z=z, theta=theta, loglik=np.array([l]), grad=g)


def bivariate_small():
    """Bivariate parsimonious Matern on 400 locations (800 rows)."""
    n = 400
    x, y, _ = datagen.generate_xy_loc(n, 2)
    e, _ = datagen.dlarnv(3, [2, 2, 2, 1], 2 * n)
    theta = np.array([1.0, 1.0, 0.1, 0.5, 1.0, 0.5])
    kern = "bivariate_matern_parsimonious"
    z = likelihood.dzvg(kern, x, y, theta, e)
    l = likelihood.dmle(kern, x, y, z, theta)
    return save("bivariate_400.npz", x=x, y=y, z=z, theta=theta, loglik=np.array([l[0]]),
                logdet=np.array([l[1]]))


def c1_csv():
    """The n = 1,600 locations and observations of c1_seed0.npz as the flat
    files the reference reads (misc/compute/flat_file.c: "x,y" per line,
    one observation per line), written with 17 significant digits so that
    they parse back to the same doubles."""
    g = np.load(os.path.join(OUT, "c1_seed0.npz"))
    locs = os.path.join(OUT, "c1_locs.csv")
    obs = os.path.join(OUT, "c1_obs.csv")
    with open(locs, "w") as f:
        for a, b in zip(g["x"], g["y"]):
            f.write(f"{a:.17g},{b:.17g}\n")
    with open(obs, "w") as f:
        for v in g["z"]:
            f.write(f"{v:.17g}\n")
    return [locs, obs]


def main():
    paths = [knu_table(), datagen_small(), c1(0), c1(1), nugget_small(), bivariate_small()] + c1_csv()
    manifest = {}
    for p in paths:
        with open(p, "rb") as f:
            manifest[os.path.basename(p)] = hashlib.sha256(f.read()).hexdigest()
    with open(os.path.join(OUT, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    main()
