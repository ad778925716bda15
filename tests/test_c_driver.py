"""A C driver built against include/exageostat_gpu.h and linked with
-lexageostat_gpu, run through the reference driver's call sequence
(tests/drivers/synthetic_dmle.c): the header is a source drop-in for a C
caller, not only for the ctypes mirror.

The CPU test compiles and links it (no compute call); the GPU test runs it
at n = 1,600, ts = 320 and checks every likelihood, the kriging
predictions and the MSE against the golden fixture c1_seed0.npz (written
by the oracle), to 1e-9 relative."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "drivers", "synthetic_dmle.c")
LIBDIR = os.path.join(ROOT, "exageostat_amd")
GOLD = os.path.join(ROOT, "tests", "golden", "c1_seed0.npz")


def _build(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libexageostat_gpu.so")):
        pytest.fail("libexageostat_gpu.so is not built")
    exe = str(tmp_path / "synthetic_dmle")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-Wno-discarded-qualifiers", "-I",
                    os.path.join(ROOT, "include"), SRC, "-L", LIBDIR, "-lexageostat_gpu",
                    f"-Wl,-rpath,{LIBDIR}", "-o", exe], check=True)
    return exe


def test_c_driver_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    assert os.path.exists(exe)
    # every undefined symbol of the driver resolves in the library
    out = subprocess.run(["ldd", exe], check=True, capture_output=True, text=True).stdout
    assert "libexageostat_gpu.so" in out and "not found" not in out


@pytest.mark.gpu
def test_c_driver_runs_the_reference_sequence(tmp_path):
    exe = _build(tmp_path)
    g = np.load(GOLD)
    thetas = "\n".join(" ".join(f"{v:.17g}" for v in th) for th in g["thetas"]) + "\n"
    print("c driver: n=1600 ts=320 seed=0 nZmiss=100", flush=True)
    r = subprocess.run([exe, "1600", "320", "0", "100", "univariate_matern_stationary"], input=thetas,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    ll = [line.split() for line in lines if line.startswith("loglik")]
    assert len(ll) == len(g["thetas"])
    # the last golden theta has cond(Sigma) = 2.2e10: one ulp of Sigma's
    # entries moves l by 1.8e-9 relative in the oracle itself (as in
    # test_gpu_parity.test_loglik_c1_golden)
    tols = [1e-9, 1e-9, 1e-9, 1e-9, 1e-7]
    for (_, l, logdet, dotp, status), ref, rld, rdp, tol in zip(ll, g["loglik"], g["logdet"], g["dotp"], tols):
        assert int(status) == 0
        assert abs(float(l) - ref) <= tol * abs(ref), (l, ref)
        assert abs(float(logdet) - rld) <= tol * abs(ref) and abs(float(dotp) - rdp) <= tol * rdp
    zmiss = np.array([float(v) for v in next(x for x in lines if x.startswith("zmiss")).split()[1:]])
    assert np.max(np.abs(zmiss - g["zmiss"])) <= 1e-9 * np.max(np.abs(g["zmiss"]))
    mse, mse_results = (float(v) for v in next(x for x in lines if x.startswith("mse")).split()[1:])
    assert abs(mse - g["mse"][0]) <= 1e-9 * g["mse"][0] and mse_results == mse
    final, iters = next(x for x in lines if x.startswith("results")).split()[1:]
    assert float(final) == float(ll[-1][1]) and int(iters) == len(ll)
