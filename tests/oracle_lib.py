"""ctypes binding of the CPU oracle (oracle/lib/liblgd_oracle.so).

Test infrastructure: used by tests/ (and by __graft_entry__.smoke / the
benchmark's CPU-baseline leg through their own loaders) as the checker.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]
ORACLE_LIB = REPO / "oracle" / "lib" / "liblgd_oracle.so"

_lib = None


def load():
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            raise FileNotFoundError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        lib = C.CDLL(str(ORACLE_LIB))
        P, L, I, D, F = C.c_void_p, C.c_long, C.c_int, C.c_double, C.c_float
        lib.lgdo_jacobi3d_f64.argtypes = [P, L, L, L, D, I, I]
        lib.lgdo_jacobi3d_f64_wrap.argtypes = [P, L, L, L, D, I, I, I]
        lib.lgdo_life2d_u8.argtypes = [P, L, L, I, I, I]
        lib.lgdo_life2d_u8_wrap.argtypes = [P, L, L, I, I, I, I]
        lib.lgdo_lbm_cavity_flags.argtypes = [P, L, L, L]
        lib.lgdo_lbm_cavity_flags.restype = None
        lib.lgdo_lbm_d3q19_double.argtypes = [P, P, L, L, L, D, D, D, D, I, I]
        lib.lgdo_lbm_d3q19_float.argtypes = [P, P, L, L, L, F, F, F, F, I, I]
        lib.lgdo_lbm_equilibrium_f64.argtypes = [P, P, P, P, P, L]
        lib.lgdo_lbm_equilibrium_f64.restype = None
        lib.lgdo_lbm_moments_f64.argtypes = [P, L, P, P, P, P]
        lib.lgdo_lbm_moments_f64.restype = None
        lib.lgdo_lbm_direction.argtypes = [I, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(I)]
        lib.lgdo_poisson_set_faces.argtypes = [P, L, P]
        lib.lgdo_poisson_set_faces.restype = None
        lib.lgdo_poisson_sweep_f32.argtypes = [P, P, P, L]
        lib.lgdo_poisson_sweep_f32.restype = None
        lib.lgdo_poisson_sweep_damped_f32.argtypes = [P, P, P, L, C.c_float]
        lib.lgdo_poisson_sweep_damped_f32.restype = None
        lib.lgdo_poisson_residual_norm.argtypes = [P, P, L]
        lib.lgdo_poisson_residual_norm.restype = D
        lib.lgdo_poisson_jacobi_f64.argtypes = [P, P, L, D, D, L, L, C.POINTER(D)]
        lib.lgdo_poisson_jacobi_f64.restype = L
        _lib = lib
    return _lib


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.c_void_p)


def jacobi3d(grid: np.ndarray, steps: int, edge: float = 0.0, threads: int = 1, wrap: int = 0) -> np.ndarray:
    """wrap: bits 1, 2, 4 make the x, y, z axes periodic (0: the Cube)."""
    out = np.ascontiguousarray(grid, dtype=np.float64).copy()
    nz, ny, nx = out.shape
    if wrap:
        rc = load().lgdo_jacobi3d_f64_wrap(_p(out), nx, ny, nz, edge, steps, threads, wrap)
    else:
        rc = load().lgdo_jacobi3d_f64(_p(out), nx, ny, nz, edge, steps, threads)
    assert rc == 0
    return out


def life2d(grid: np.ndarray, steps: int, edge: int = 0, threads: int = 1, wrap: int = 0) -> np.ndarray:
    """wrap: bits 1, 2 make the x, y axes periodic (0: the Cube)."""
    out = np.ascontiguousarray(grid, dtype=np.uint8).copy()
    ny, nx = out.shape
    if wrap:
        rc = load().lgdo_life2d_u8_wrap(_p(out), nx, ny, edge, steps, threads, wrap)
    else:
        rc = load().lgdo_life2d_u8(_p(out), nx, ny, edge, steps, threads)
    assert rc == 0
    return out


def lbm_directions():
    lib = load()
    out = []
    for i in range(19):
        ex, ey, ez, op = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib.lgdo_lbm_direction(i, C.byref(ex), C.byref(ey), C.byref(ez), C.byref(op))
        out.append((ex.value, ey.value, ez.value, op.value))
    return out


def lbm_cavity_flags(nx: int, ny: int, nz: int) -> np.ndarray:
    fl = np.zeros((nz, ny, nx), dtype=np.uint8)
    load().lgdo_lbm_cavity_flags(_p(fl), nx, ny, nz)
    return fl


def lbm_run(f: np.ndarray, flags: np.ndarray, steps: int, omega: float, lid=(0.0, 0.0, 0.0),
            threads: int = 1) -> np.ndarray:
    out = np.ascontiguousarray(f).copy()
    _, nz, ny, nx = out.shape
    lib = load()
    if out.dtype == np.float64:
        rc = lib.lgdo_lbm_d3q19_double(_p(out), _p(flags), nx, ny, nz, omega, *lid, steps, threads)
    else:
        assert out.dtype == np.float32
        rc = lib.lgdo_lbm_d3q19_float(_p(out), _p(flags), nx, ny, nz, omega, *lid, steps, threads)
    assert rc == 0
    return out


def lbm_equilibrium(rho, jx, jy, jz) -> np.ndarray:
    rho, jx, jy, jz = (np.ascontiguousarray(a, dtype=np.float64) for a in (rho, jx, jy, jz))
    f = np.empty((19,) + rho.shape, dtype=np.float64)
    load().lgdo_lbm_equilibrium_f64(_p(f), _p(rho), _p(jx), _p(jy), _p(jz), rho.size)
    return f


def lbm_moments(f: np.ndarray):
    f = np.ascontiguousarray(f, dtype=np.float64)
    shp = f.shape[1:]
    out = [np.empty(shp, dtype=np.float64) for _ in range(4)]
    load().lgdo_lbm_moments_f64(_p(f), int(np.prod(shp)), *(_p(o) for o in out))
    return tuple(out)


def poisson_faces(u: np.ndarray, faces) -> None:
    fa = np.ascontiguousarray(faces, dtype=np.float64)
    load().lgdo_poisson_set_faces(_p(u), u.shape[0] - 2, _p(fa))


def poisson_sweep_f32(e: np.ndarray, r: np.ndarray) -> np.ndarray:
    out = np.zeros_like(e)
    load().lgdo_poisson_sweep_f32(_p(e), _p(out), _p(r), e.shape[0] - 2)
    return out


def poisson_sweep_damped_f32(e: np.ndarray, r: np.ndarray, omega: float) -> np.ndarray:
    out = np.zeros_like(e)
    load().lgdo_poisson_sweep_damped_f32(_p(e), _p(out), _p(r), e.shape[0] - 2, omega)
    return out


def poisson_residual_norm(u: np.ndarray, f: np.ndarray) -> float:
    return float(load().lgdo_poisson_residual_norm(_p(u), _p(f), u.shape[0] - 2))


def poisson_jacobi(u: np.ndarray, f: np.ndarray, rel_tol: float, reference: float = 0.0, max_sweeps: int = 10 ** 7,
                   check: int = 1):
    """Plain FP64 Jacobi in place; returns (sweeps, final residual norm)."""
    fn = C.c_double(0.0)
    sweeps = load().lgdo_poisson_jacobi_f64(_p(u), _p(f), u.shape[0] - 2, rel_tol, reference, max_sweeps, check,
                                            C.byref(fn))
    return int(sweeps), fn.value


def _conv_lib():
    lib = load()
    if not getattr(lib, "_conv_ready", False):
        P, L, I = C.c_void_p, C.c_long, C.c_int
        lib.lgdo_conv3d_forward.argtypes = [P, L, I, I, P, P, I, P]
        lib.lgdo_conv3d_forward.restype = None
        lib.lgdo_conv3d_dgrad.argtypes = [P, L, I, I, P, I, P]
        lib.lgdo_conv3d_dgrad.restype = None
        lib.lgdo_conv3d_wgrad.argtypes = [P, L, I, I, P, I, P, P]
        lib.lgdo_conv3d_wgrad.restype = None
        lib._conv_ready = True
    return lib


def conv3d_forward(x: np.ndarray, w: np.ndarray, bias=None, ksize: int = 3) -> np.ndarray:
    """x[n,n,n,cin], w[cout,taps,cin] -> y[n,n,n,cout] in FP64."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    n, cin, cout = x.shape[0], x.shape[-1], w.shape[0]
    y = np.zeros(x.shape[:3] + (cout,))
    b = None if bias is None else np.ascontiguousarray(bias, dtype=np.float64)
    _conv_lib().lgdo_conv3d_forward(_p(x), n, cin, ksize, _p(w), None if b is None else _p(b), cout, _p(y))
    return y


def conv3d_dgrad(dy: np.ndarray, w: np.ndarray, ksize: int = 3) -> np.ndarray:
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    n, cout, cin = dy.shape[0], w.shape[0], w.shape[-1]
    dx = np.zeros(dy.shape[:3] + (cin,))
    _conv_lib().lgdo_conv3d_dgrad(_p(dy), n, cin, ksize, _p(w), cout, _p(dx))
    return dx


def conv3d_wgrad(x: np.ndarray, dy: np.ndarray, ksize: int = 3):
    x = np.ascontiguousarray(x, dtype=np.float64)
    dy = np.ascontiguousarray(dy, dtype=np.float64)
    n, cin, cout = x.shape[0], x.shape[-1], dy.shape[-1]
    taps = 1 if ksize == 1 else 27
    dw = np.zeros((cout, taps, cin))
    db = np.zeros(cout)
    _conv_lib().lgdo_conv3d_wgrad(_p(x), n, cin, ksize, _p(dy), cout, _p(dw), _p(db))
    return dw, db
