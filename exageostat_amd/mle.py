"""Python mirror of the reference's driver-level interface for the exact
likelihood path (examples/synthetic_dmle_test.c, src/compute/MLE.c).

Every method calls the C ABI of libexageostat_gpu.so; nothing here computes
covariances, factorisations or likelihoods itself.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from ._lib import MLEData, Desc, c_double_p

KERNEL_PARAMS = {
    "univariate_matern_stationary": 3,
    "univariate_matern_nuggets_stationary": 4,
    "bivariate_matern_parsimonious": 6,
    "bivariate_matern_parsimonious_profile": 6,
    "bivariate_matern_parsimonious2": 6,
    "bivariate_matern_parsimonious2_profile": 6,
    "univariate_spacetime_matern_stationary": 7,
}


def _dp(a: np.ndarray):
    return a.ctypes.data_as(c_double_p)


def _check(lib, what: str):
    st = lib.exageostat_gpu_last_status()
    if st < 0:
        raise RuntimeError(f"{what} failed ({st}): {lib.exageostat_gpu_last_error().decode()}")
    return st


def device_count() -> int:
    return _lib.load().exageostat_gpu_device_count()


def generate_xy_loc(n: int, seed: int):
    """GenerateXYLoc through the library (glibc srand/rand + Morton sort)."""
    lib = _lib.load()
    loc = lib.GenerateXYLoc(n, seed)
    x = np.ctypeslib.as_array(loc.contents.x, shape=(n,)).copy()
    y = np.ctypeslib.as_array(loc.contents.y, shape=(n,)).copy()
    libc = ctypes.CDLL("libc.so.6")
    libc.free.argtypes = [ctypes.c_void_p]
    libc.free(ctypes.cast(loc.contents.x, ctypes.c_void_p))
    libc.free(ctypes.cast(loc.contents.y, ctypes.c_void_p))
    libc.free(ctypes.cast(loc, ctypes.c_void_p))
    return x, y


def generate_xy_loc_st(n: int, t_slots: int, seed: int):
    """GenerateXYLoc_ST through the library: (x, y, t), n points per slot."""
    lib = _lib.load()
    loc = lib.GenerateXYLoc_ST(n, t_slots, seed)
    total = n * t_slots
    out = [np.ctypeslib.as_array(getattr(loc.contents, f), shape=(total,)).copy() for f in ("x", "y", "z")]
    libc = ctypes.CDLL("libc.so.6")
    libc.free.argtypes = [ctypes.c_void_p]
    for f in ("x", "y", "z"):
        libc.free(ctypes.cast(getattr(loc.contents, f), ctypes.c_void_p))
    libc.free(ctypes.cast(loc, ctypes.c_void_p))
    return tuple(out)


def dlarnv(idist: int, iseed, n: int):
    """LAPACK dlarnv through the library; returns (values, advanced iseed)."""
    lib = _lib.load()
    seed = (ctypes.c_int * 4)(*[int(v) for v in iseed])
    out = np.empty(n)
    lib.exageostat_dlarnv(idist, seed, n, _dp(out))
    return out, list(seed)


def synchronize() -> None:
    _lib.load().exageostat_gpu_synchronize()


def profile_enable(on: bool = True) -> None:
    lib = _lib.load()
    lib.exageostat_gpu_profile_reset()
    lib.exageostat_gpu_profile_enable(1 if on else 0)


def profile_get(name: str) -> dict:
    lib = _lib.load()
    launches = ctypes.c_long(0)
    ms = ctypes.c_double(0)
    fl = ctypes.c_double(0)
    by = ctypes.c_double(0)
    lib.exageostat_gpu_profile_get(name.encode(), ctypes.byref(launches), ctypes.byref(ms), ctypes.byref(fl),
                                   ctypes.byref(by))
    return {"launches": launches.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}


class ExactMLE:
    """One MLE_data with its GPU context: EXAGEOSTAT_dmle_Call at
    construction, MLE_Finalize at close().

    t: third coordinate of every location (time for the space-time kernel),
    or None.  computation: "exact" or "diag_approx" (the band-limited
    factorisation, band = diag_thick tiles).  lean: no FP64 covariance on
    the device (MLE_data.mp_lean); only precision 1 and 2 likelihoods,
    prediction and MLOE/MMOM run, and z must come from set_z."""

    def __init__(self, x, y, ts: int, kernel: str = "univariate_matern_stationary", dm: str = "ed",
                 precision: int = 0, silent: bool = True, t=None, computation: str = "exact",
                 diag_thick: int = 1, lean: bool = False):
        self.lib = _lib.load()
        if kernel not in KERNEL_PARAMS:
            raise ValueError(f"unsupported kernel {kernel!r}")
        self.kernel = kernel
        self.nparams = KERNEL_PARAMS[kernel]
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.t = None if t is None else np.ascontiguousarray(t, dtype=np.float64)
        self.nloc = len(self.x)
        self.N = self.nloc * (2 if kernel.startswith("bivariate") else 1)
        self.ts = ts
        self._strings = [computation.encode(), kernel.encode(), dm.encode(), b""]
        d = MLEData()
        d.computation = self._strings[0]
        d.kernel_fun = self._strings[1]
        d.dm = self._strings[2]
        d.recovery_file = self._strings[3]
        d.precision = precision
        d.silent = 1 if silent else 0
        d.l1.x = _dp(self.x)
        d.l1.y = _dp(self.y)
        d.l1.z = _dp(self.t) if self.t is not None else None
        d.diag_thick = diag_thick
        d.mp_ir_steps = 3
        d.mp_lean = 1 if lean else 0
        self.data = d
        zero = ctypes.c_int(0)
        one = ctypes.c_int(1)
        tsi = ctypes.c_int(ts)
        self.lib.exageostat_init(ctypes.byref(one), ctypes.byref(one), ctypes.byref(tsi), ctypes.byref(zero))
        _check(self.lib, "exageostat_init")
        self.lib.EXAGEOSTAT_dmle_Call(ctypes.byref(d), 1, 1, ts, 1, 1, self.N, self.N, 0)
        _check(self.lib, "EXAGEOSTAT_dmle_Call")
        self._pred_refs = []

    # -- data ------------------------------------------------------------
    def set_z(self, z) -> None:
        """Observations (the real-data path: Lapack_to_Tile into descZ)."""
        z = np.ascontiguousarray(z, dtype=np.float64)
        assert z.shape == (self.N,)
        desc = ctypes.cast(self.data.descZ, ctypes.POINTER(Desc))
        self.lib.EXAGEOSTAT_Lapack_to_Tile(_dp(z), self.N, desc)
        _check(self.lib, "Lapack_to_Tile")
        self.data.iter_count = 0

    def zvg(self, theta, e) -> None:
        """MLE_zvg: z = L(theta) e on the GPU."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        e = np.ascontiguousarray(e, dtype=np.float64)
        self.lib.MLE_zvg(ctypes.byref(self.data), _dp(e), _dp(th), self.N, self.ts, 0, 1, 1)
        _check(self.lib, "MLE_zvg")

    def get_z(self) -> np.ndarray:
        z = np.empty(self.N)
        self.lib.MLE_get_zobs(ctypes.byref(self.data), _dp(z), self.N)
        _check(self.lib, "MLE_get_zobs")
        return z

    # -- likelihood ------------------------------------------------------
    def loglik(self, theta, grad: bool = False):
        """MLE_alg; returns l (and d l/d theta when grad=True)."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        g = np.zeros(len(th)) if grad else None
        val = self.lib.MLE_alg(len(th), _dp(th), _dp(g) if grad else None, ctypes.byref(self.data))
        if self.lib.exageostat_gpu_last_status() != -3:
            # -3: the mixed refinement did not reach mp_ir_tol; l is still
            # returned and self.status says so
            _check(self.lib, "MLE_alg")
        return (val, g) if grad else val

    def set_mixed(self, fp32_corr: float = 0.0, bf16: bool = False, fp64_corr: float = 0.0,
                  diag_thick: int = 1, ir_steps: int = 3, lolo: bool = False, ir_tol: float = 0.0,
                  grad_f32: int = 0) -> None:
        """Mixed-precision options (used when precision == 2): tiles whose
        largest correlation is >= fp64_corr stay FP64, tiles below fp32_corr
        get BF16 updates when bf16 is on, the rest are FP32; the solve is
        refined at most ir_steps times in FP64, stopping early once a step
        changes z^T Sigma^-1 z by at most ir_tol (relative; 0: always
        ir_steps steps).  lolo: the FP32 tiles' split-BF16 products keep
        their lo*lo term (four products instead of three).  grad_f32: 1 (or
        True) runs a gradient request's Takahashi sweep with its column
        stage on the FP32 matrix cores (faster; its error grows with the
        condition number of Sigma) and reports the swept inverse's residual
        along z in data.mp_grad_check; 2 also recomputes the gradient with
        the FP64 sweep when that residual exceeds 1e-6."""
        d = self.data
        d.mp_fp32_corr = fp32_corr
        d.mp_bf16 = 1 if bf16 else 0
        d.mp_fp64_corr = fp64_corr
        d.diag_thick = diag_thick
        d.mp_ir_steps = ir_steps
        d.mp_lolo = 1 if lolo else 0
        d.mp_ir_tol = ir_tol
        d.mp_grad_f32 = int(grad_f32)

    @property
    def precision(self) -> int:
        return self.data.precision

    @precision.setter
    def precision(self, p: int) -> None:
        self.data.precision = p

    def prec_map(self) -> np.ndarray:
        """Per-tile precision classes (0 FP64, 1 FP32, 3 BF16 updates) of the
        last mixed evaluation as an mt x mt lower-triangular array (-1 above
        the diagonal)."""
        ntiles = self.lib.exageostat_mixed_prec_map(ctypes.byref(self.data), None, 0)
        _check(self.lib, "exageostat_mixed_prec_map")
        buf = (ctypes.c_ubyte * max(ntiles, 1))()
        self.lib.exageostat_mixed_prec_map(ctypes.byref(self.data), buf, ntiles)
        mt = (self.N + self.ts - 1) // self.ts
        out = -np.ones((mt, mt), dtype=np.int8)
        idx = 0
        for n in range(mt):
            for m in range(n, mt):
                out[m, n] = buf[idx] if ntiles else -1
                idx += 1
        return out

    @property
    def status(self) -> int:
        return self.data.status

    @property
    def logdet(self) -> float:
        return self.data.logdet

    @property
    def variances(self):
        """(variance1, variance2) of the last profile-kernel evaluation."""
        return self.data.variance1, self.data.variance2

    @property
    def dotp(self) -> float:
        return self.data.dotp

    def phase_times(self) -> dict:
        d = self.data
        return {"gen": d.time_gen, "facto": d.time_facto, "logdet": d.time_logdet, "solve": d.time_solve,
                "total": d.time_total}

    # -- Fisher information ----------------------------------------------
    def fisher(self, theta):
        """Fisher information F (np x np) and its inverse at theta
        (exageostat_fisher; EXAGEOSTAT_Fisher_Tile returns the inverse)."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        npar = len(th)
        F = np.zeros((npar, npar))
        Finv = np.zeros((npar, npar))
        st = self.lib.exageostat_fisher(ctypes.byref(self.data), _dp(th), _dp(F), _dp(Finv))
        if st == -1:
            _check(self.lib, "exageostat_fisher")
        # st > 0: Sigma(theta) not positive definite (data.status); st == -2:
        # F not positive definite, Finv is NaN
        return F, Finv

    # -- prediction ------------------------------------------------------
    def predict(self, theta, xobs, yobs, zobs, xmiss, ymiss, zactual=None, tobs=None, tmiss=None, precision=None):
        """Kriging; returns (zmiss, mse).  tobs/tmiss: third coordinates
        (time) of the observed and held-out locations.  precision (default:
        the model's): 0 EXAGEOSTAT_dmle_Predict_Tile (FP64), 1
        EXAGEOSTAT_smle_Predict_Tile (C22 factored with FP32 off-diagonal
        tiles), 2 EXAGEOSTAT_sdmle_Predict_Tile (C22 factored under the
        set_mixed options, the solve refined in FP64)."""
        xo = np.ascontiguousarray(xobs, dtype=np.float64)
        yo = np.ascontiguousarray(yobs, dtype=np.float64)
        xm = np.ascontiguousarray(xmiss, dtype=np.float64)
        ym = np.ascontiguousarray(ymiss, dtype=np.float64)
        zo = np.ascontiguousarray(zobs, dtype=np.float64)
        rows = 2 if self.kernel.startswith("bivariate") else 1
        n_miss = len(xm) * rows
        n_obs = len(xo) * rows
        zm = np.zeros(n_miss)
        za = None if zactual is None else np.ascontiguousarray(zactual, dtype=np.float64)
        to = None if tobs is None else np.ascontiguousarray(tobs, dtype=np.float64)
        tm = None if tmiss is None else np.ascontiguousarray(tmiss, dtype=np.float64)
        self._pred_refs = [xo, yo, xm, ym, to, tm]
        d = self.data
        d.lobs.x, d.lobs.y, d.lobs.z = _dp(xo), _dp(yo), (_dp(to) if to is not None else None)
        d.lmiss.x, d.lmiss.y, d.lmiss.z = _dp(xm), _dp(ym), (_dp(tm) if tm is not None else None)
        th = np.ascontiguousarray(theta, dtype=np.float64)
        self.lib.EXAGEOSTAT_dmle_Predict_Allocate(ctypes.byref(d), len(xm), len(xo), self.ts, 1, 1, 1)
        _check(self.lib, "Predict_Allocate")
        prec = self.precision if precision is None else precision
        name = {0: "EXAGEOSTAT_dmle_Predict_Tile", 1: "EXAGEOSTAT_smle_Predict_Tile",
                2: "EXAGEOSTAT_sdmle_Predict_Tile"}[prec]
        mse = getattr(self.lib, name)(ctypes.byref(d), _dp(th), n_miss, n_obs, _dp(zo),
                                      _dp(za) if za is not None else None, _dp(zm), self.N)
        if self.lib.exageostat_gpu_last_status() != -3:
            # -3: the mixed refinement of the solve did not reach mp_ir_tol;
            # the predictions are still returned and self.status says so
            _check(self.lib, name)
        return zm, mse

    def mloe_mmom(self, theta_true, theta_est, xobs, yobs, xmiss, ymiss):
        """EXAGEOSTAT_dmle_mloe_mmom_Tile through exageostat_mloe_mmom;
        returns (mloe, mmom, loe, mom) with the per-location values."""
        xo = np.ascontiguousarray(xobs, dtype=np.float64)
        yo = np.ascontiguousarray(yobs, dtype=np.float64)
        xm = np.ascontiguousarray(xmiss, dtype=np.float64)
        ym = np.ascontiguousarray(ymiss, dtype=np.float64)
        self._mloe_refs = [xo, yo, xm, ym]
        d = self.data
        d.lobs.x, d.lobs.y, d.lobs.z = _dp(xo), _dp(yo), None
        d.lmiss.x, d.lmiss.y, d.lmiss.z = _dp(xm), _dp(ym), None
        tt = np.ascontiguousarray(theta_true, dtype=np.float64)
        te = np.ascontiguousarray(theta_est, dtype=np.float64)
        loe = np.zeros(len(xm))
        mom = np.zeros(len(xm))
        self.lib.mloe_mmom_init(ctypes.byref(d), len(xm), len(xo), self.ts, 1, 1)
        _check(self.lib, "mloe_mmom_init")
        rc = self.lib.exageostat_mloe_mmom(ctypes.byref(d), _dp(tt), _dp(te), len(xm), len(xo), _dp(loe), _dp(mom))
        if rc < 0:
            _check(self.lib, "exageostat_mloe_mmom")
        return d.mloe, d.mmom, loe, mom

    def close(self) -> None:
        if self.data is not None and self.data.gpu:
            self.lib.prediction_finalize(ctypes.byref(self.data))
            self.lib.mloe_mmom_finalize(ctypes.byref(self.data))
            self.lib.MLE_Finalize(ctypes.byref(self.data))
        self.data = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMatrix:
    """A descriptor-level handle (EXAGEOSTAT_Desc_Create) for stage tests:
    dcmg -> dpotrf -> dmdet -> dtrsv, the reference's call sequence."""

    def __init__(self, m: int, n: int, ts: int, lower: bool = True):
        self.lib = _lib.load()
        self.desc = ctypes.POINTER(Desc)()
        st = self.lib.EXAGEOSTAT_Desc_Create(ctypes.byref(self.desc), _lib.ChamLower if lower else
                                             _lib.ChamUpperLower, ts, m, n)
        if st != 0:
            _check(self.lib, "Desc_Create")
            raise RuntimeError("Desc_Create failed")
        self.m, self.n, self.ts, self.lower = m, n, ts, lower
        self._keep = []

    def dcmg(self, theta, x1, y1, x2=None, y2=None, kernel="univariate_matern_stationary", dm="ed", t1=None,
             t2=None):
        """EXAGEOSTAT_MLE_dcmg_Tile_Async; t1/t2: third coordinates (time
        for the space-time kernel) of the row and column locations."""
        x1 = np.ascontiguousarray(x1, dtype=np.float64)
        y1 = np.ascontiguousarray(y1, dtype=np.float64)
        x2 = x1 if x2 is None else np.ascontiguousarray(x2, dtype=np.float64)
        y2 = y1 if y2 is None else np.ascontiguousarray(y2, dtype=np.float64)
        t1 = None if t1 is None else np.ascontiguousarray(t1, dtype=np.float64)
        t2 = t1 if t2 is None else np.ascontiguousarray(t2, dtype=np.float64)
        l1 = _lib.Location(_dp(x1), _dp(y1), _dp(t1) if t1 is not None else None)
        l2 = _lib.Location(_dp(x2), _dp(y2), _dp(t2) if t2 is not None else None)
        th = np.ascontiguousarray(theta, dtype=np.float64)
        self._keep = [x1, y1, x2, y2, th, t1, t2]
        st = self.lib.EXAGEOSTAT_MLE_dcmg_Tile_Async(_lib.ChamLower, self.desc, ctypes.byref(l1), ctypes.byref(l2),
                                                     None, _dp(th), dm.encode(), kernel.encode(), None, None)
        _check(self.lib, "dcmg")
        return st

    def potrf(self) -> int:
        info = self.lib.EXAGEOSTAT_dpotrf_Tile(_lib.ChamLower, self.desc)
        _check(self.lib, "dpotrf")
        return info

    def potri(self) -> None:
        """Sigma^-1 in place of the factor (call right after potrf)."""
        self.lib.EXAGEOSTAT_dpotri_Tile(_lib.ChamLower, self.desc)
        _check(self.lib, "dpotri")

    def logdet_half(self) -> float:
        acc = np.zeros(1)
        ud = ctypes.POINTER(Desc)()
        self.lib.EXAGEOSTAT_Desc_Create_User(ctypes.byref(ud), _dp(acc), 1, 1)
        self.lib.EXAGEOSTAT_MLE_dmdet_Tile_Async(self.desc, None, None, ud)
        _check(self.lib, "dmdet")
        self.lib.EXAGEOSTAT_Desc_Destroy(ctypes.byref(ud))
        return float(acc[0])

    def trsv(self, b, trans: bool = False) -> np.ndarray:
        b = np.array(b, dtype=np.float64, copy=True)
        self.lib.EXAGEOSTAT_dtrsv_Tile(_lib.ChamTrans if trans else _lib.ChamNoTrans, self.desc, _dp(b))
        _check(self.lib, "dtrsv")
        return b

    def to_numpy(self) -> np.ndarray:
        A = np.zeros((self.m, self.n), order="F")
        self.lib.EXAGEOSTAT_Tile_to_Lapack(self.desc, A.ctypes.data_as(c_double_p), self.m)
        _check(self.lib, "Tile_to_Lapack")
        return A

    def from_numpy(self, A) -> None:
        A = np.asfortranarray(A, dtype=np.float64)
        self.lib.EXAGEOSTAT_Lapack_to_Tile(A.ctypes.data_as(c_double_p), self.m, self.desc)
        _check(self.lib, "Lapack_to_Tile")

    def close(self):
        if self.desc:
            self.lib.EXAGEOSTAT_Desc_Destroy(ctypes.byref(self.desc))
            self.desc = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def loglik_formula(dotp: float, logdet: float, n: int) -> float:
    """l = -dotp/2 - logdet/2 - (n/2) log 2 pi (MLE_dexact.c:426)."""
    return -0.5 * dotp - 0.5 * logdet - (n / 2.0) * math.log(2.0 * math.pi)
