"""ctypes binding of libqms.so (declared in include/qms.h).

There is no CPU fallback: if the library is missing or no GPU is visible, the
calls raise.  Arrays cross the boundary as numpy float64 / uint16 buffers.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqms.so")

QMS_OK = 0
QMS_ERR_ARG = -1
QMS_ERR_HIP = -2
QMS_ERR_NOCONV = -3
QMS_ERR_NOMEM = -4
QMS_ERR_STATE = -5
QMS_BUSY = 1
QMS_FP64 = 0
QMS_MIXED = 1

# Every symbol include/qms.h declares.
EXPORTS = (
    "qms_config_default", "qms_create", "qms_steady", "qms_step", "qms_run",
    "qms_get_fields", "qms_set_fields", "qms_get_h", "qms_get_coeffs", "qms_apply", "qms_residual", "qms_rhs",
    "qms_get_stats", "qms_set_zchunk", "qms_set_colwidth", "qms_use_coef_classes", "qms_time_kernel", "qms_gram_schmidt", "qms_get_work", "qms_destroy", "qms_last_error",
    "qms_run_async", "qms_poll", "qms_wait", "qms_estimate_bytes",
    "qms_batch_create", "qms_batch_set_blocks", "qms_batch_apply", "qms_batch_tune", "qms_batch_time",
    "qms_batch_plan", "qms_batch_last_error", "qms_batch_destroy", "qms_gram_schmidt_noalloc",
)

# qms_time_kernel selectors
K_APPLY, K_SMOOTH_A, K_SMOOTH_V, K_COPY, K_RHS, K_HFIELD, K_RESID, K_GRAM_SCHMIDT = range(8)
K_VCYCLE_A, K_VCYCLE_V, K_ITERATION, K_COPY_SUSTAINED = 8, 9, 10, 11
# or-ed into a selector: the kernel timed with cold caches
K_COLD = 256


class QmsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libqms error {code}: {msg}")
        self.code = code


class NotConverged(QmsError):
    pass


class Config(ctypes.Structure):
    _fields_ = [
        ("cells", ctypes.c_int * 3),
        ("h", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("bc", ctypes.c_double * 12),
        ("precision", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("restart", ctypes.c_int),
        ("max_iters", ctypes.c_int),
        ("mg_nu", ctypes.c_int),
        ("mg_omega", ctypes.c_double),
        ("mg_alpha", ctypes.c_double),
        ("use_graphs", ctypes.c_int),
        ("inner_tol", ctypes.c_double),
        ("mg_nu_coarse", ctypes.c_int),
        ("floor_tol", ctypes.c_double),
        ("mg_fused_tail", ctypes.c_int),
        ("mg_omega_v", ctypes.c_double),
        ("mg_nu_v", ctypes.c_int),
        ("mg_alpha_v", ctypes.c_double),
        ("mg_cheb_lmax", ctypes.c_double),
        ("mg_cheb_ratio", ctypes.c_double),
        ("mg_cheb_lmax_v", ctypes.c_double),
        ("mg_cheb_ratio_v", ctypes.c_double),
        ("mg_fused_restrict", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("last_iters", ctypes.c_int),
        ("last_restarts", ctypes.c_int),
        ("last_relres", ctypes.c_double),
        ("last_relres0", ctypes.c_double),
        ("total_iters", ctypes.c_int64),
        ("solves", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
        ("n_levels_a", ctypes.c_int),
        ("n_levels_v", ctypes.c_int),
        ("coef_classes", ctypes.c_int),
        ("queued_iters", ctypes.c_int64),
        ("mg_classes_a", ctypes.c_int),
        ("mg_classes_v", ctypes.c_int),
        ("stalled_solves", ctypes.c_int64),
        ("fused_restrict_blocks", ctypes.c_int),
    ]


_lib = None


def load(path: Optional[str] = None):
    """Load libqms.so; raises if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} is missing: build it with `python -m alpineqms_amd.build`")
    L = ctypes.CDLL(p)
    d = ctypes.POINTER(ctypes.c_double)
    H = ctypes.c_void_p
    L.qms_config_default.argtypes = [ctypes.POINTER(Config)]
    L.qms_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_uint16), d, ctypes.c_int,
                             ctypes.POINTER(ctypes.c_void_p)]
    L.qms_steady.argtypes = [H, d]
    L.qms_step.argtypes = [H, d]
    L.qms_run.argtypes = [H, ctypes.c_int, d]
    L.qms_run_async.argtypes = [H, ctypes.c_int, d]
    L.qms_poll.argtypes = [H]
    L.qms_wait.argtypes = [H]
    L.qms_estimate_bytes.argtypes = [ctypes.POINTER(Config)]
    L.qms_estimate_bytes.restype = ctypes.c_int64
    L.qms_get_fields.argtypes = [H, d]
    L.qms_set_fields.argtypes = [H, d]
    L.qms_get_h.argtypes = [H, ctypes.c_int, d]
    L.qms_get_coeffs.argtypes = [H, d]
    L.qms_apply.argtypes = [H, d, ctypes.c_double, d, d]
    L.qms_rhs.argtypes = [H, d, ctypes.c_double, d, d]
    L.qms_get_stats.argtypes = [H, ctypes.POINTER(Stats)]
    L.qms_set_zchunk.argtypes = [H, ctypes.c_int]
    L.qms_time_kernel.argtypes = [H, ctypes.c_int, d, ctypes.c_int, d, d]
    L.qms_get_work.argtypes = [H, d]
    L.qms_gram_schmidt.argtypes = [H, ctypes.c_int, d, d, d, d]
    L.qms_residual.argtypes = [H, d, ctypes.c_double, d, d, d]
    L.qms_gram_schmidt_noalloc.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p]
    L.qms_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.qms_batch_set_blocks.argtypes = [H, ctypes.c_int]
    L.qms_batch_apply.argtypes = [H, d, ctypes.c_int]
    L.qms_batch_tune.argtypes = [H, d, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.qms_set_colwidth.argtypes = [H, ctypes.c_int]
    L.qms_use_coef_classes.argtypes = [H, ctypes.c_int]
    L.qms_batch_time.argtypes = [H, ctypes.c_int, d, ctypes.c_int, ctypes.c_int, d, d]
    ip = ctypes.POINTER(ctypes.c_int)
    up = ctypes.POINTER(ctypes.c_uint)
    L.qms_batch_plan.argtypes = [ctypes.c_int, ip, ip, ip, ctypes.c_int, up, ctypes.c_int, up, ctypes.c_int, ip, ip]
    L.qms_batch_last_error.argtypes = [H]
    L.qms_batch_last_error.restype = ctypes.c_char_p
    L.qms_batch_destroy.argtypes = [H]
    L.qms_batch_destroy.restype = None
    L.qms_destroy.argtypes = [H]
    L.qms_destroy.restype = None
    L.qms_last_error.argtypes = [H]
    L.qms_last_error.restype = ctypes.c_char_p
    for name in EXPORTS:
        if name not in ("qms_destroy", "qms_last_error", "qms_batch_destroy", "qms_batch_last_error",
                        "qms_estimate_bytes"):
            getattr(L, name).restype = ctypes.c_int
    if path is None:
        _lib = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_config() -> Config:
    c = Config()
    load().qms_config_default(ctypes.byref(c))
    return c


def make_config(cells: Sequence[int], h: Sequence[float], dt: float, bc: Sequence[float],
                precision: int = QMS_FP64, tol: float = 1e-12, restart: int = 30, max_iters: int = 3000,
                mg_nu: int = 4, mg_omega: float = 0.7, mg_alpha: float = 1.5, use_graphs: bool = True,
                inner_tol: float = 1e-6, mg_nu_coarse: int = 3, floor_tol: float = 0.0,
                mg_fused_tail: bool = True, mg_omega_v: float = 0.85, mg_nu_v: int = 6,
                mg_alpha_v: float = 1.9, mg_cheb_lmax: float = 2.6, mg_cheb_ratio: float = 4.0,
                mg_cheb_lmax_v: float = 2.0, mg_cheb_ratio_v: float = 8.0,
                mg_fused_restrict: bool = True) -> Config:
    cfg = default_config()
    for a in range(3):
        cfg.cells[a] = int(cells[a])
        cfg.h[a] = float(h[a])
    cfg.dt = float(dt)
    bcv = np.asarray(bc, dtype=np.float64).ravel()
    if bcv.size != 12:
        raise ValueError("bc needs 12 values (B0 and the 3x3 gradient)")
    for a in range(12):
        cfg.bc[a] = float(bcv[a])
    cfg.precision = int(precision)
    cfg.tol = float(tol)
    cfg.restart = int(restart)
    cfg.max_iters = int(max_iters)
    cfg.mg_nu = int(mg_nu)
    cfg.mg_omega = float(mg_omega)
    cfg.mg_alpha = float(mg_alpha)
    cfg.use_graphs = 1 if use_graphs else 0
    cfg.inner_tol = float(inner_tol)
    cfg.mg_nu_coarse = int(mg_nu_coarse)
    cfg.floor_tol = float(floor_tol)
    cfg.mg_fused_tail = 1 if mg_fused_tail else 0
    cfg.mg_omega_v = float(mg_omega_v)
    cfg.mg_nu_v = int(mg_nu_v)
    cfg.mg_alpha_v = float(mg_alpha_v)
    cfg.mg_cheb_lmax = float(mg_cheb_lmax)
    cfg.mg_cheb_ratio = float(mg_cheb_ratio)
    cfg.mg_cheb_lmax_v = float(mg_cheb_lmax_v)
    cfg.mg_cheb_ratio_v = float(mg_cheb_ratio_v)
    cfg.mg_fused_restrict = 1 if mg_fused_restrict else 0
    return cfg


# qms_gram_schmidt_noalloc limits (include/qms.h)
NOALLOC_MAX_VECTORS = 64
NOALLOC_SCRATCH_BYTES = 1 << 20


def gram_schmidt_noalloc(n: int, nv: int, basis_dev: int, w_dev: int, hcoef_dev: int, stream: int = 0) -> None:
    """qms_gram_schmidt_noalloc on device addresses the caller owns (queued on
    `stream`, a hipStream_t as an integer; 0 is the null stream).  Allocates
    nothing; raises QmsError on a refused argument or launch."""
    L = load()
    rc = L.qms_gram_schmidt_noalloc(int(n), int(nv), ctypes.c_void_p(basis_dev), ctypes.c_void_p(w_dev),
                                    ctypes.c_void_p(hcoef_dev), ctypes.c_void_p(stream))
    if rc != QMS_OK:
        raise QmsError(rc, L.qms_last_error(None).decode())


def estimate_bytes(cells: Sequence[int], precision: int = QMS_FP64, restart: int = 30) -> int:
    """Device bytes a handle of this size holds once it has stepped (host-only,
    qms_estimate_bytes): the admission measure of the serving loop."""
    cfg = make_config(cells, (1.0, 1.0, 1.0), 1.0, np.zeros(12), precision=precision, restart=restart)
    return int(load().qms_estimate_bytes(ctypes.byref(cfg)))


class Solver:
    """One design on the GPU (a ``qms_handle``)."""

    def __init__(self, cells: Sequence[int], h: Sequence[float], dt: float, bc: Sequence[float],
                 cell_slot: np.ndarray, slot_props: np.ndarray, precision: int = QMS_FP64,
                 tol: float = 1e-12, restart: int = 30, max_iters: int = 3000, mg_nu: int = 4,
                 mg_omega: float = 0.7, mg_alpha: float = 1.5, use_graphs: bool = True, inner_tol: float = 1e-6,
                 mg_nu_coarse: int = 3, floor_tol: float = 0.0, mg_fused_tail: bool = True,
                 mg_omega_v: float = 0.85, mg_nu_v: int = 6, mg_alpha_v: float = 1.9,
                 mg_cheb_lmax: float = 2.6, mg_cheb_ratio: float = 4.0, mg_cheb_lmax_v: float = 2.0,
                 mg_cheb_ratio_v: float = 8.0, mg_fused_restrict: bool = True):
        L = load()
        cfg = make_config(cells, h, dt, bc, precision, tol, restart, max_iters, mg_nu, mg_omega, mg_alpha,
                          use_graphs, inner_tol, mg_nu_coarse, floor_tol, mg_fused_tail, mg_omega_v, mg_nu_v,
                          mg_alpha_v, mg_cheb_lmax, mg_cheb_ratio, mg_cheb_lmax_v, mg_cheb_ratio_v,
                          mg_fused_restrict)
        self.cells = tuple(int(c) for c in cells)
        self.nodes = tuple(c + 1 for c in self.cells)
        self.dt = float(dt)
        slot = np.ascontiguousarray(cell_slot, dtype=np.uint16).ravel()
        if slot.size != np.prod(self.cells):
            raise ValueError("cell_slot must hold cells_x*cells_y*cells_z entries")
        props = np.ascontiguousarray(slot_props, dtype=np.float64).reshape(-1, 6)
        out = ctypes.c_void_p()
        rc = L.qms_create(ctypes.byref(cfg), slot.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), _dp(props),
                          props.shape[0], ctypes.byref(out))
        if rc != QMS_OK:
            raise QmsError(rc, L.qms_last_error(None).decode())
        self._h = out
        self._L = L
        self.precision = precision

    # -- helpers -----------------------------------------------------------
    def _check(self, rc: int):
        if rc != QMS_OK:
            msg = self._L.qms_last_error(self._h).decode()
            if rc == QMS_ERR_NOCONV:
                raise NotConverged(rc, msg)
            raise QmsError(rc, msg)

    @property
    def n1(self) -> int:
        return int(np.prod(self.nodes))

    def close(self):
        if getattr(self, "_h", None):
            self._L.qms_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- solver entry points -------------------------------------------------
    def steady(self, v):
        vv = np.ascontiguousarray(v, dtype=np.float64)
        self._check(self._L.qms_steady(self._h, _dp(vv)))

    def step(self, v):
        vv = np.ascontiguousarray(v, dtype=np.float64)
        self._check(self._L.qms_step(self._h, _dp(vv)))

    def run(self, velocities: np.ndarray):
        vv = np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1, 3)
        self._check(self._L.qms_run(self._h, vv.shape[0], _dp(vv)))

    def run_async(self, velocities: np.ndarray):
        """Start the steps on a library thread (returns at once)."""
        vv = np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1, 3)
        self._check(self._L.qms_run_async(self._h, vv.shape[0], _dp(vv)))

    def poll(self) -> bool:
        """True while an asynchronous run is going; raises its error if it failed."""
        rc = self._L.qms_poll(self._h)
        if rc == QMS_BUSY:
            return True
        self._check(rc)
        return False

    def wait(self):
        self._check(self._L.qms_wait(self._h))

    def get_fields(self) -> np.ndarray:
        nx, ny, nz = self.nodes
        out = np.empty((4, nz, ny, nx), dtype=np.float64)
        self._check(self._L.qms_get_fields(self._h, _dp(out)))
        return out

    def set_fields(self, f: np.ndarray):
        nx, ny, nz = self.nodes
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(4, nz, ny, nx)
        self._check(self._L.qms_set_fields(self._h, _dp(a)))

    def get_h(self, unit_mu: bool = False) -> np.ndarray:
        cx, cy, cz = self.cells
        out = np.empty((3, cz, cy, cx), dtype=np.float64)
        self._check(self._L.qms_get_h(self._h, 1 if unit_mu else 0, _dp(out)))
        return out

    def get_coeffs(self) -> np.ndarray:
        """(6, Nz, Ny, Nx): Mu_x, Mu_y, Mu_z, Sigma_x, Sigma_y, Sigma_z."""
        nx, ny, nz = self.nodes
        out = np.empty((6, nz, ny, nx), dtype=np.float64)
        self._check(self._L.qms_get_coeffs(self._h, _dp(out)))
        return out

    def apply(self, v, dt: float, x: np.ndarray) -> np.ndarray:
        vv = np.ascontiguousarray(v, dtype=np.float64)
        xx = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(xx)
        self._check(self._L.qms_apply(self._h, _dp(vv), float(dt), _dp(xx), _dp(y)))
        return y

    def residual(self, v, dt: float, x: np.ndarray, b: np.ndarray) -> np.ndarray:
        """D^-1 (b - M x) on free rows, b - x on fixed rows."""
        vv = np.ascontiguousarray(v, dtype=np.float64)
        xx = np.ascontiguousarray(x, dtype=np.float64)
        bb = np.ascontiguousarray(b, dtype=np.float64)
        r = np.empty_like(xx)
        self._check(self._L.qms_residual(self._h, _dp(vv), float(dt), _dp(xx), _dp(bb), _dp(r)))
        return r

    def rhs(self, v, dt: float, aprev: np.ndarray) -> np.ndarray:
        vv = np.ascontiguousarray(v, dtype=np.float64)
        a = np.ascontiguousarray(aprev, dtype=np.float64)
        b = np.empty_like(a)
        self._check(self._L.qms_rhs(self._h, _dp(vv), float(dt), _dp(a), _dp(b)))
        return b

    def stats(self) -> Stats:
        s = Stats()
        self._check(self._L.qms_get_stats(self._h, ctypes.byref(s)))
        return s

    def set_zchunk(self, planes: int):
        """Planes per operator workgroup (z-march depth of the LDS-staged
        operator kernel)."""
        self._check(self._L.qms_set_zchunk(self._h, int(planes)))

    def use_coef_classes(self, on: bool):
        """Operator coefficients from the class table (True) or streamed."""
        self._check(self._L.qms_use_coef_classes(self._h, 1 if on else 0))

    def set_colwidth(self, nodes: int):
        """Operator column width in nodes (32, the default, or 64)."""
        self._check(self._L.qms_set_colwidth(self._h, int(nodes)))

    def time_kernel(self, which: int, reps: int, v=None):
        """Mean milliseconds and algorithmic bytes per launch of one kernel
        (selectors K_*; or-ed with K_COLD, with cold caches), timed with HIP
        events on the handle's stream."""
        ms = ctypes.c_double()
        by = ctypes.c_double()
        vp = None
        if v is not None:
            vv = np.ascontiguousarray(v, dtype=np.float64)
            vp = _dp(vv)
        self._check(self._L.qms_time_kernel(self._h, which, vp, reps, ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    def gram_schmidt(self, basis: np.ndarray, w: np.ndarray):
        """One Gram-Schmidt step of the solver's FGMRES on the GPU: returns
        (h, w') with h[i] = V_i . w for the nv basis vectors, w' = w - sum h_i
        V_i and h[nv] = w' . w'."""
        nv = basis.shape[0]
        bb = np.ascontiguousarray(basis, dtype=np.float64)
        ww = np.ascontiguousarray(w, dtype=np.float64)
        hc = np.empty(nv + 1, dtype=np.float64)
        out = np.empty_like(ww)
        self._check(self._L.qms_gram_schmidt(self._h, int(nv), _dp(bb), _dp(ww), _dp(hc), _dp(out)))
        return hc, out

    def get_work(self) -> np.ndarray:
        """The last operator output D^-1 M x (4 node-sized arrays)."""
        nx, ny, nz = self.nodes
        out = np.empty((4, nz, ny, nx), dtype=np.float64)
        self._check(self._L.qms_get_work(self._h, _dp(out)))
        return out


def batch_plan(ncols: Sequence[int], nz: Sequence[int], band: Sequence[int], blocks: int):
    """Schedule of a batched launch, computed on the host (no GPU call).

    Returns (segments, slot_off): segments is an (S, 4) array of (design,
    column, first plane, end plane); the workgroup of slot s runs rows
    slot_off[s]:slot_off[s+1] in order."""
    L = load()
    n = len(ncols)
    a = [(ctypes.c_int * n)(*map(int, x)) for x in (ncols, nz, band)]
    ns, nsl = ctypes.c_int(), ctypes.c_int()
    rc = L.qms_batch_plan(n, a[0], a[1], a[2], int(blocks), None, 0, None, 0, ctypes.byref(ns), ctypes.byref(nsl))
    if rc != QMS_OK:
        raise QmsError(rc, L.qms_last_error(None).decode())
    segs = np.zeros((ns.value, 4), dtype=np.uint32)
    off = np.zeros(nsl.value + 1, dtype=np.uint32)
    up = ctypes.POINTER(ctypes.c_uint)
    rc = L.qms_batch_plan(n, a[0], a[1], a[2], int(blocks), segs.ctypes.data_as(up), ns.value,
                          off.ctypes.data_as(up), nsl.value + 1, ctypes.byref(ns), ctypes.byref(nsl))
    if rc != QMS_OK:
        raise QmsError(rc, L.qms_last_error(None).decode())
    return segs, off


class Batch:
    """Many designs in one persistent operator launch (a ``qms_batch``).
    The solvers must stay alive and idle while the batch is used."""

    def __init__(self, solvers: Sequence[Solver]):
        L = load()
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s._h for s in self.solvers])
        out = ctypes.c_void_p()
        rc = L.qms_batch_create(arr, len(self.solvers), ctypes.byref(out))
        if rc != QMS_OK:
            raise QmsError(rc, L.qms_batch_last_error(None).decode())
        self._b = out
        self._L = L

    def _check(self, rc: int):
        if rc != QMS_OK:
            raise QmsError(rc, self._L.qms_batch_last_error(self._b).decode())

    def _v(self, v):
        vv = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (len(self.solvers), 3)))
        return vv

    def set_blocks(self, blocks: int):
        self._check(self._L.qms_batch_set_blocks(self._b, int(blocks)))

    def apply(self, v, transient: bool = True):
        """Operator output of every design into its solver's work vector."""
        vv = self._v(v)
        self._check(self._L.qms_batch_apply(self._b, _dp(vv), 1 if transient else 0))

    def tune(self, v, transient: bool = True, reps: int = 10):
        vv = self._v(v)
        kz = (ctypes.c_int * len(self.solvers))()
        cw = (ctypes.c_int * len(self.solvers))()
        bl = ctypes.c_int()
        self._check(self._L.qms_batch_tune(self._b, _dp(vv), 1 if transient else 0, int(reps), kz, cw,
                                           ctypes.byref(bl)))
        return list(kz), list(cw), bl.value

    def time(self, v, transient: bool = True, reps: int = 20, sequential: bool = False):
        vv = self._v(v)
        ms = ctypes.c_double()
        by = ctypes.c_double()
        self._check(self._L.qms_batch_time(self._b, 1 if sequential else 0, _dp(vv), 1 if transient else 0,
                                           int(reps), ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    def close(self):
        if getattr(self, "_b", None):
            self._L.qms_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
