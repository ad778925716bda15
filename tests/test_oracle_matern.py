"""Pins of the oracle's Matern kernels and of the library's own K_nu (host
build of the same header the GPU kernels use)."""
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest
import scipy.special as sp

from oracle import datagen, matern

GOLD = os.path.join(os.path.dirname(__file__), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_scipy_kv_against_mpmath_table():
    """The oracle's K_nu (scipy kv) within 1e-13 relative of 30-digit mpmath."""
    g = np.load(os.path.join(GOLD, "knu_table.npz"))
    k = sp.kv(g["nu"], g["x"])
    rel = np.abs(k - g["k"]) / np.abs(g["k"])
    assert np.max(rel) < 1e-13


def test_mpmath_table_regenerates():
    import mpmath as mp
    g = np.load(os.path.join(GOLD, "knu_table.npz"))
    mp.mp.dps = 30
    for i in range(0, len(g["nu"]), 17):
        assert float(mp.besselk(g["nu"][i], g["x"][i])) == pytest.approx(g["k"][i], rel=1e-15)


def test_half_integer_closed_forms():
    x, y, _ = datagen.generate_xy_loc(200, 1)
    d = matern.distance_matrix(x, y, x, y)
    s2, beta = 1.7, 0.13
    c05 = matern.core_dcmg(x, y, x, y, (s2, beta, 0.5))
    c15 = matern.core_dcmg(x, y, x, y, (s2, beta, 1.5))
    r = d / beta
    assert np.max(np.abs(c05 - s2 * np.exp(-r))) < 1e-14
    assert np.max(np.abs(c15 - s2 * (1 + r) * np.exp(-r))) < 1e-14


def test_dcmg_basic_properties():
    x, y, _ = datagen.generate_xy_loc(300, 2)
    th = (1.3, 0.1, 1.2)
    S = matern.core_dcmg(x, y, x, y, th)
    assert np.allclose(np.diag(S), th[0], rtol=0, atol=0)
    assert np.array_equal(S, S.T)
    assert np.min(np.linalg.eigvalsh(S)) > 0
    Sn = matern.core_dcmg_nuggets(x, y, x, y, th + (0.2,))
    assert np.allclose(np.diag(Sn), th[0] + 0.2, atol=0)
    off = ~np.eye(300, dtype=bool)
    assert np.array_equal(Sn[off], S[off])


def test_bivariate_reduces_to_univariate():
    """nu1 = nu2, sigma1 = sigma2 and rho = 1 make all four blocks equal."""
    x, y, _ = datagen.generate_xy_loc(100, 3)
    nu = 0.8
    B = matern.core_dcmg_bivariate_parsimonious(x, y, x, y, (1.4, 1.4, 0.1, nu, nu, 1.0))
    U = matern.core_dcmg(x, y, x, y, (1.4, 0.1, nu))
    for a in (0, 1):
        for b in (0, 1):
            assert np.max(np.abs(B[a::2, b::2] - U)) < 1e-14


def test_bivariate_is_spd_and_cross_constant():
    x, y, _ = datagen.generate_xy_loc(150, 4)
    th = (1.0, 2.0, 0.1, 0.5, 1.0, 0.5)
    B = matern.core_dcmg_bivariate_parsimonious(x, y, x, y, th)
    assert np.min(np.linalg.eigvalsh(B)) > 0
    nu12, rho_t, s12 = matern.parsimonious_constants(th)
    expected = 0.5 * math.sqrt(math.gamma(1.5) * math.gamma(2.0) / (math.gamma(0.5) * math.gamma(1.0))) \
        * math.gamma(0.75) / math.gamma(1.75)
    assert rho_t == pytest.approx(expected, rel=1e-15)
    assert B[1, 0] == pytest.approx(s12, rel=1e-15)


@pytest.mark.parametrize("theta", [(1.0, 0.1, 0.5), (0.7, 0.2, 1.37), (1.5, 0.05, 2.2)])
def test_derivative_kernels_match_finite_differences(theta):
    x, y, _ = datagen.generate_xy_loc(80, 5)
    dS = matern.matern_derivatives(x, y, x, y, theta, nugget=True)
    base = np.array(theta + (0.03,))
    for p in range(4):
        h = 1e-6 * max(1.0, abs(base[p]))
        tp = base.copy()
        tm = base.copy()
        tp[p] += h
        tm[p] -= h
        fd = (matern.core_dcmg_nuggets(x, y, x, y, tp) - matern.core_dcmg_nuggets(x, y, x, y, tm)) / (2 * h)
        assert np.max(np.abs(dS[p] - fd)) < 2e-7 * max(1.0, np.max(np.abs(fd)))


@pytest.mark.parametrize("theta", [(1.0, 1.3, 0.1, 0.5, 1.0, 0.5), (0.8, 1.1, 0.2, 1.37, 0.6, -0.3)])
def test_bivariate_derivatives_match_finite_differences(theta):
    """Each of the six analytic derivative matrices of the parsimonious
    bivariate kernel against central differences of the kernel itself."""
    x, y, _ = datagen.generate_xy_loc(60, 5)
    dS = matern.bivariate_derivatives(x, y, x, y, theta)
    base = np.array(theta)
    for p in range(6):
        h = 1e-6 * max(1.0, abs(base[p]))
        tp, tm = base.copy(), base.copy()
        tp[p] += h
        tm[p] -= h
        fd = (matern.core_dcmg_bivariate_parsimonious(x, y, x, y, tp)
              - matern.core_dcmg_bivariate_parsimonious(x, y, x, y, tm)) / (2 * h)
        assert np.max(np.abs(dS[p] - fd)) < 2e-7 * max(1.0, np.max(np.abs(fd))), p


def test_dkv_dnu_against_mpmath_table():
    g = np.load(os.path.join(GOLD, "knu_table.npz"))
    sel = g["x"] >= 1e-3
    d = np.array([matern.dkv_dnu(nu, np.array([x]))[0] for nu, x in zip(g["nu"][sel], g["x"][sel])])
    ref = g["dk_dnu"][sel]
    assert np.max(np.abs(d - ref) / np.maximum(np.abs(ref), 1e-300)) < 1e-13


_KNU_DRIVER = r"""
#include "besselk.hpp"
#include <cstdio>
#include <cstdlib>
int main(int argc, char** argv) {
  FILE* f = fopen(argv[1], "r");
  double nu, x;
  while (fscanf(f, "%lf %lf", &nu, &x) == 2) {
    exa::KnuOrder<double> o; exa::knu_prepare(nu, o);
    exa::KnuOrder<exa::Dual> od; exa::knu_prepare(nu, od);
    double k, k1; exa::knu_pair(o, x, k, k1);
    exa::Dual kd, kd1; exa::knu_pair(od, x, kd, kd1);
    printf("%.17g %.17g %.17g\n", k, k1, kd.d);
  }
  return 0;
}
"""


def test_library_knu_host_build_against_mpmath():
    """The K_nu the GPU kernels use (csrc/besselk.hpp, compiled for the host
    with hipcc): K_nu and K_{nu+1} within 1e-14 relative of 30-digit mpmath
    (worst case measured 6.2e-15 at nu = 0.1, x = 1.7, the Temme branch near
    its x = 2 limit), dK_nu/dnu (dual numbers) within 5e-12."""
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    g = np.load(os.path.join(GOLD, "knu_table.npz"))
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "knu.cpp")
        exe = os.path.join(td, "knu")
        inp = os.path.join(td, "in.txt")
        open(src, "w").write(_KNU_DRIVER)
        np.savetxt(inp, np.stack([g["nu"], g["x"]], axis=1), fmt="%.17g")
        subprocess.run([hipcc, "-O2", "-std=c++17", "--offload-arch=gfx950", "-x", "hip",
                        "-I", os.path.join(ROOT, "exageostat_amd", "csrc"), src, "-o", exe],
                       check=True, capture_output=True)
        out = subprocess.run([exe, inp], check=True, capture_output=True, text=True).stdout
    vals = np.array([[float(v) for v in line.split()] for line in out.strip().splitlines()])
    assert np.max(np.abs(vals[:, 0] - g["k"]) / g["k"]) < 1e-14
    assert np.max(np.abs(vals[:, 1] - g["k1"]) / g["k1"]) < 1e-14
    assert np.max(np.abs(vals[:, 2] - g["dk_dnu"]) / np.abs(g["dk_dnu"])) < 5e-12
