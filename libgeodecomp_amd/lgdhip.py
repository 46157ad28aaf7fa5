"""ctypes binding of the liblgdhip C-ABI (include/lgdhip.h).

This is the Python side of the drop-in boundary: tests and the benchmark
drive the same `extern "C"` entry points that the C++ HipSimulator calls.
The library is loaded from the package's `lib/` directory (built by
`libgeodecomp_amd.build`); if it is missing, loading raises -- there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "liblgdhip.so"


class LgdHipError(RuntimeError):
    """A liblgdhip entry point returned a non-zero status."""

    def __init__(self, code: int, func: str, detail: str):
        super().__init__(f"{func} failed with {code}: {detail}")
        self.code = code


LAYOUT_MEMBER_PLANES, LAYOUT_Z_SLABS = 0, 1
GRID_VIEW = 0x1


class lgdhip_grid(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("halo", C.c_int32), ("elem_bytes", C.c_int32), ("members", C.c_int32),
        ("layout", C.c_int32), ("flags", C.c_int32),
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("ox", C.c_int64), ("oy", C.c_int64), ("oz", C.c_int64),
        ("sx", C.c_int64), ("sy", C.c_int64), ("sz", C.c_int64),
    ]


class lgdhip_cell_grid(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("halo", C.c_int32), ("cell_bytes", C.c_int64),
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("px", C.c_int64), ("py", C.c_int64), ("pz", C.c_int64),
    ]


class lgdhip_tile(C.Structure):
    _fields_ = [
        ("bx", C.c_int32), ("by", C.c_int32), ("bz", C.c_int32),
        ("steps_per_launch", C.c_int32), ("variant", C.c_int32), ("flags", C.c_int32),
    ]


class lgdhip_lbm_params(C.Structure):
    _fields_ = [("omega", C.c_float), ("lid_ux", C.c_float), ("lid_uy", C.c_float), ("lid_uz", C.c_float)]


class lgdhip_poisson_problem(C.Structure):
    _fields_ = [("n", C.c_int32), ("reserved", C.c_int32), ("offset", C.c_int64), ("face", C.c_double * 6)]


class lgdhip_poisson_options(C.Structure):
    _fields_ = [("rel_tol", C.c_double), ("chunk", C.c_int32), ("max_sweeps", C.c_int32),
                ("inner_reduction", C.c_double), ("use_graph", C.c_int32), ("sweep_variant", C.c_int32),
                ("reference_norms", C.POINTER(C.c_double)), ("smooth_sweeps", C.c_int32),
                ("smooth_omega", C.c_float), ("final_fp64_sweeps", C.c_int32)]


class lgdhip_poisson_result(C.Structure):
    _fields_ = [("sweeps", C.c_int64), ("outer_iterations", C.c_int32), ("converged", C.c_int32),
                ("reference_norm", C.c_double), ("residual_norm", C.c_double)]


LBM_WALL_XLO, LBM_WALL_XHI = 0x01, 0x02
LBM_WALL_YLO, LBM_WALL_YHI = 0x04, 0x08
LBM_WALL_ZLO, LBM_WALL_ZHI = 0x10, 0x20
LBM_LID_YHI = 0x40

_P = C.c_void_p
_I = C.c_int
_I64 = C.c_int64
_SZ = C.c_size_t
_GP = C.POINTER(lgdhip_grid)
_CGP = C.POINTER(lgdhip_cell_grid)
_TP = C.POINTER(lgdhip_tile)
_IP = C.POINTER(C.c_int)

# name -> (restype, argtypes); every symbol declared in include/lgdhip.h
SIGNATURES = {
    "lgdhip_version": (C.c_char_p, []),
    "lgdhip_error_string": (C.c_char_p, [_I]),
    "lgdhip_last_error": (C.c_char_p, []),
    "lgdhip_get_device_count": (_I, [_IP]),
    "lgdhip_set_device": (_I, [_I]),
    "lgdhip_device_name": (_I, [_I, C.c_char_p, _I]),
    "lgdhip_device_synchronize": (_I, []),
    "lgdhip_malloc": (_I, [C.POINTER(_P), _SZ]),
    "lgdhip_free": (_I, [_P]),
    "lgdhip_host_malloc": (_I, [C.POINTER(_P), _SZ]),
    "lgdhip_host_free": (_I, [_P]),
    "lgdhip_memcpy_h2d": (_I, [_P, _P, _SZ, _P]),
    "lgdhip_memcpy_d2h": (_I, [_P, _P, _SZ, _P]),
    "lgdhip_memcpy_d2d": (_I, [_P, _P, _SZ, _P]),
    "lgdhip_memcpy2d_d2h": (_I, [_P, _SZ, _P, _SZ, _SZ, _SZ, _P]),
    "lgdhip_get_device": (_I, [_IP]),
    "lgdhip_fill_u8": (_I, [_P, C.c_uint8, _SZ, _P]),
    "lgdhip_copy_kernel": (_I, [_P, _P, _SZ, _P]),
    "lgdhip_stream_create": (_I, [C.POINTER(_P)]),
    "lgdhip_stream_destroy": (_I, [_P]),
    "lgdhip_stream_synchronize": (_I, [_P]),
    "lgdhip_event_create": (_I, [C.POINTER(_P)]),
    "lgdhip_event_destroy": (_I, [_P]),
    "lgdhip_event_record": (_I, [_P, _P]),
    "lgdhip_stream_wait_event": (_I, [_P, _P]),
    "lgdhip_event_synchronize": (_I, [_P]),
    "lgdhip_event_elapsed_ms": (_I, [C.POINTER(C.c_float), _P, _P]),
    "lgdhip_graph_begin": (_I, [_P]),
    "lgdhip_graph_end": (_I, [_P, C.POINTER(_P)]),
    "lgdhip_graph_launch": (_I, [_P, _P]),
    "lgdhip_graph_destroy": (_I, [_P]),
    "lgdhip_grid_make": (_I, [_GP, _I, _I64, _I64, _I64, _I, _I, _I]),
    "lgdhip_grid_set_layout": (_I, [_GP, _I]),
    "lgdhip_grid_view": (_I, [_GP, _I64, _I64, _I64, _I64, _I64, _I64, _GP]),
    "lgdhip_grid_plane_elems": (_I64, [_GP]),
    "lgdhip_grid_bytes": (_I64, [_GP]),
    "lgdhip_grid_fill_edge": (_I, [_P, _GP, _P, _P]),
    "lgdhip_grid_fill_halo": (_I, [_P, _GP, _P, _I, _P]),
    "lgdhip_cell_grid_make": (_I, [_CGP, _I, _I64, _I64, _I64, _I, _I64]),
    "lgdhip_cell_grid_bytes": (_I64, [_CGP]),
    "lgdhip_cell_grid_fill_halo": (_I, [_P, _CGP, _P, _I, _P]),
    "lgdhip_cell_grid_load": (_I, [_P, _CGP, _P, _P]),
    "lgdhip_cell_grid_store": (_I, [_P, _CGP, _P, _P]),
    "lgdhip_grid_load": (_I, [_P, _GP, _P, _P]),
    "lgdhip_grid_store": (_I, [_P, _GP, _P, _P]),
    "lgdhip_jacobi3d_f64_run": (_I, [_P, _P, _GP, _I, _TP, _IP, _P]),
    "lgdhip_jacobi3d_f64_plan": (_I, [_GP, _TP, _IP, _IP]),
    "lgdhip_jacobi3d_f64_run_zrange": (_I, [_P, _P, _GP, _I, _I, _TP, _IP, _P]),
    "lgdhip_jacobi3d_f64_run_periodic": (_I, [_P, _P, _GP, _I, _I, C.c_double, _TP, _IP, _P]),
    "lgdhip_life2d_u8_run": (_I, [_P, _P, _GP, _I, _IP, _P]),
    "lgdhip_life2d_packed_words": (_I64, [_GP]),
    "lgdhip_life2d_pack": (_I, [_P, _GP, _P, _P]),
    "lgdhip_life2d_unpack": (_I, [_P, _GP, _P, _P]),
    "lgdhip_life2d_packed_run": (_I, [_P, _P, _GP, _I, _I, _TP, _IP, _P]),
    "lgdhip_life2d_packed_run_periodic": (_I, [_P, _P, _GP, _I, _I, _I, _TP, _IP, _P]),
    "lgdhip_life2d_workspace_bytes": (_SZ, [_GP]),
    "lgdhip_life2d_run": (_I, [_P, _GP, _I, _I, _P, _SZ, _TP, _P]),
    "lgdhip_lbm_d3q19_f32_run": (_I, [_P, _P, _P, _GP, C.POINTER(lgdhip_lbm_params), _I, _TP, _IP, _P]),
    "lgdhip_lbm_cavity_flags": (_I, [_P, _GP, _P]),
    "lgdhip_lbm_d3q19_f32_moments": (_I, [_P, _GP, _P, _P, _P, _P, _P]),
    "lgdhip_poisson_layout": (_I64, [C.POINTER(lgdhip_poisson_problem), _I]),
    "lgdhip_poisson_sweep_plan": (_I, [C.POINTER(lgdhip_poisson_problem), _I, _I, _IP]),
    "lgdhip_poisson_workspace_bytes": (_SZ, [C.POINTER(lgdhip_poisson_problem), _I]),
    "lgdhip_poisson_set_faces": (_I, [_P, C.POINTER(lgdhip_poisson_problem), _I, _P, _SZ, _P]),
    "lgdhip_poisson_residual_norms": (_I, [_P, _P, C.POINTER(lgdhip_poisson_problem), _I, _P, _SZ,
                                           C.POINTER(C.c_double), _P]),
    "lgdhip_poisson_sweeps_f32": (_I, [_P, _P, _P, C.POINTER(lgdhip_poisson_problem), _I, _I, _TP, _P, _SZ, _IP,
                                       _P]),
    "lgdhip_poisson_sweeps_f64": (_I, [_P, _P, _P, C.POINTER(lgdhip_poisson_problem), _I, _I, _P, _SZ, _IP, _P]),
    "lgdhip_poisson_sweeps_f32_noalloc": (_I, [_P, _P, _P, C.POINTER(lgdhip_poisson_problem), _I, _I, _TP, _IP,
                                               _P]),
    "lgdhip_poisson_damped_sweeps_f32": (_I, [_P, _P, _P, C.POINTER(lgdhip_poisson_problem), _I, _I, C.c_float, _TP,
                                              _P, _SZ, _IP, _P]),
    "lgdhip_poisson_batch_solve": (_I, [_P, _P, C.POINTER(lgdhip_poisson_problem), _I,
                                        C.POINTER(lgdhip_poisson_options), _P, _SZ,
                                        C.POINTER(lgdhip_poisson_result), _P]),
    "lgdhip_conv3d_kpad": (_I, [_I, _I]),
    "lgdhip_conv3d_bf16_forward": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I, _I, _P, _I, _I, _I, _P]),
    "lgdhip_conv3d_forward_workspace_bytes": (_SZ, [_I, _I, _I, _I]),
    "lgdhip_conv3d_bf16_forward_ws": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I, _I, _P, _I, _I, _I, _P, _SZ,
                                           _P]),
    "lgdhip_conv3d_dgrad_weights": (_I, [_P, _I, _I, _I, _P, _P]),
    "lgdhip_conv3d_wgrad_workspace_bytes": (_SZ, [_I, _I, _I, _I, _I]),
    "lgdhip_conv3d_bf16_wgrad": (_I, [_P, _I, _I, _I, _I, _I, _P, _I, _I, _I, _P, _P, _I, _I, _P, _SZ, _P]),
    "lgdhip_relu_backward_bf16": (_I, [_P, _I, _I, _P, _I, _I, _I, _I64, _P]),
    "lgdhip_avgpool2_bf16": (_I, [_P, _I, _I, _I, _P, _I, _I, _I, _P]),
    "lgdhip_avgpool2_backward_bf16": (_I, [_P, _I, _I, _I, _P, _I, _I, _I, _I, _P]),
    "lgdhip_upsample2_bf16": (_I, [_P, _I, _I, _I, _P, _I, _I, _I, _P]),
    "lgdhip_upsample2_backward_bf16": (_I, [_P, _I, _I, _I, _P, _I, _I, _I, _I, _P]),
    "lgdhip_upsample2_concat_bf16": (_I, [_P, _I, _I, _I, _I, _P, _I, _I, _I, _P, _I, _I, _P]),
    "lgdhip_pack_bf16": (_I, [_P, _I64, _P, _I, _I, _P]),
    "lgdhip_mse_partials": (_I64, [_I64]),
    "lgdhip_mse_loss": (_I, [_P, _P, _I64, _P, _I, _P, _P]),
    "lgdhip_adam_step": (_I, [_P, _P, _P, _P, _I64, _I, _I, C.c_float, C.c_float, C.c_float, C.c_float, _I,
                              _P, _P]),
    "lgdhip_poisson_unet_input": (_I, [_P, C.POINTER(lgdhip_poisson_problem), C.c_double, _I, _P, _P, _I, _P]),
    "lgdhip_poisson_unet_target": (_I, [_P, C.POINTER(lgdhip_poisson_problem), C.c_double, _I, _P, _P, _P]),
    "lgdhip_poisson_unet_guess": (_I, [_P, C.POINTER(lgdhip_poisson_problem), C.c_double, _P, _P, _P]),
    "lgdhip_poisson_coarse_spacing": (C.c_double, [_I, _I]),
    "lgdhip_poisson_coarse_restrict": (_I, [_P, C.POINTER(lgdhip_poisson_problem), _P,
                                            C.POINTER(lgdhip_poisson_problem), _P]),
    "lgdhip_poisson_coarse_prolong": (_I, [_P, C.POINTER(lgdhip_poisson_problem), _I, _P, _P]),
}

# lgdhip_conv3d_kpad returns a value, not a status
_VALUE_RETURNING_INT = {"lgdhip_conv3d_kpad"}

# functions returning a status code (checked); the others return values
_STATUS = {n for n, (r, _) in SIGNATURES.items() if r is _I} - _VALUE_RETURNING_INT

_lib = None


def library_path() -> Path:
    return Path(os.environ.get("LGDHIP_LIBRARY", str(LIB_PATH)))


def load(path: Path | None = None):
    """Load liblgdhip.so and attach the signatures.  Raises if it is absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else library_path()
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code: int, func: str) -> None:
    if code != 0:
        lib = load()
        detail = lib.lgdhip_last_error().decode() or lib.lgdhip_error_string(code).decode()
        raise LgdHipError(code, func, detail)


def call(name: str, *args):
    """Call an entry point; raise LgdHipError on a non-zero status."""
    lib = load()
    r = getattr(lib, name)(*args)
    if name in _STATUS:
        check(r, name)
    return r


# ----------------------------------------------------------------------
# small RAII-style helpers used by tests and the benchmark
# ----------------------------------------------------------------------
def device_count() -> int:
    n = C.c_int(0)
    call("lgdhip_get_device_count", C.byref(n))
    return n.value


def device_name(dev: int = 0) -> str:
    buf = C.create_string_buffer(256)
    call("lgdhip_device_name", dev, buf, 256)
    return buf.value.decode()


class DeviceBuffer:
    """Owning device allocation."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        call("lgdhip_malloc", C.byref(self.ptr), self.nbytes)

    @property
    def value(self) -> int:
        return self.ptr.value or 0

    def free(self):
        if self.ptr:
            call("lgdhip_free", self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def upload(self, arr: np.ndarray, stream=None):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        call("lgdhip_memcpy_h2d", self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, stream)

    def download(self, arr: np.ndarray, stream=None):
        assert arr.flags.c_contiguous and arr.nbytes <= self.nbytes
        call("lgdhip_memcpy_d2h", arr.ctypes.data_as(C.c_void_p), self.ptr, arr.nbytes, stream)


def make_grid(dim: int, nx: int, ny: int, nz: int = 1, halo: int = 1, elem_bytes: int = 8,
              members: int = 1, layout: int = LAYOUT_MEMBER_PLANES) -> lgdhip_grid:
    g = lgdhip_grid()
    call("lgdhip_grid_make", C.byref(g), dim, nx, ny, nz, halo, elem_bytes, members)
    if layout != LAYOUT_MEMBER_PLANES:
        call("lgdhip_grid_set_layout", C.byref(g), layout)
    return g


def grid_bytes(g: lgdhip_grid) -> int:
    return int(load().lgdhip_grid_bytes(C.byref(g)))


class DeviceGrid:
    """A padded device SoA grid plus its descriptor."""

    def __init__(self, g: lgdhip_grid, dtype):
        self.g = g
        self.dtype = np.dtype(dtype)
        assert self.dtype.itemsize == g.elem_bytes
        self.buf = DeviceBuffer(grid_bytes(g))

    @property
    def ptr(self):
        return self.buf.ptr

    def logical_shape(self):
        g = self.g
        if g.dim == 2:
            shp = (g.ny, g.nx)
        else:
            shp = (g.nz, g.ny, g.nx)
        return ((g.members,) + shp) if g.members > 1 else shp

    def fill_edge(self, edge, stream=None):
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(edge, dtype=self.dtype), (self.g.members,)))
        call("lgdhip_grid_fill_edge", self.ptr, C.byref(self.g), e.ctypes.data_as(C.c_void_p), stream)

    def load(self, host: np.ndarray, stream=None):
        h = np.ascontiguousarray(host, dtype=self.dtype)
        assert h.shape == self.logical_shape(), (h.shape, self.logical_shape())
        call("lgdhip_grid_load", self.ptr, C.byref(self.g), h.ctypes.data_as(C.c_void_p), stream)

    def store(self, stream=None) -> np.ndarray:
        out = np.empty(self.logical_shape(), dtype=self.dtype)
        call("lgdhip_grid_store", self.ptr, C.byref(self.g), out.ctypes.data_as(C.c_void_p), stream)
        return out


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        call("lgdhip_stream_create", C.byref(self.h))

    def sync(self):
        call("lgdhip_stream_synchronize", self.h)

    def __del__(self):
        try:
            if self.h:
                call("lgdhip_stream_destroy", self.h)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        call("lgdhip_event_create", C.byref(self.h))

    def record(self, stream=None):
        call("lgdhip_event_record", self.h, stream.h if isinstance(stream, Stream) else stream)

    def sync(self):
        call("lgdhip_event_synchronize", self.h)

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float(0.0)
        call("lgdhip_event_elapsed_ms", C.byref(ms), self.h, later.h)
        return float(ms.value)

    def __del__(self):
        try:
            if self.h:
                call("lgdhip_event_destroy", self.h)
        except Exception:
            pass


TILE_XCD_SWIZZLE = 0x1
TILE_NO_XCD_SWIZZLE = 0x2
TILE_XCD_ROWS = 0x4


def tile(bx=0, by=0, bz=0, steps_per_launch=0, variant=0, flags=0) -> lgdhip_tile:
    return lgdhip_tile(bx, by, bz, steps_per_launch, variant, flags)
