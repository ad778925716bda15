"""CPU oracle for the exact-dense Gaussian log-likelihood path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ``exageostat_amd``
package, its HIP library or its C-ABI) may import, link or call anything
under ``oracle/``.  Only ``tests/``, the repository entry point's
``smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it, and only
as the checker.

It restates, in numpy/scipy (floating point) and plain Python/C (integer
and RNG work), what the reference computes on this path:

* ``datagen``  -- GenerateXYLoc / zsort_locations / dlarnv / shuffle /
  locations_obs_zsort_inplace / pick_random_points
  (misc/compute/MLE_misc.c, examples/synthetic_dmle_test.c).
* ``matern``   -- core_dcmg, core_dcmg_nuggets,
  core_dcmg_bivariate_parsimonious and the Matern derivative kernels
  (exageostat_exact/cpu_core/compute/core_zcmg.c, core_zcmg_nuggets.c).
* ``likelihood`` -- EXAGEOSTAT_dmle_Tile, EXAGEOSTAT_MLE_dzvg_Tile,
  EXAGEOSTAT_dmle_Predict_Tile, the dense adjoint gradient,
  EXAGEOSTAT_Fisher_Tile and EXAGEOSTAT_dmle_mloe_mmom_Tile
  (exageostat_exact/src/compute/MLE_dexact.c).

How each function is pinned (golden vectors, library cross-checks,
closed forms) is stated in the docstring of that function and exercised by
``tests/test_oracle_*.py``.
"""
