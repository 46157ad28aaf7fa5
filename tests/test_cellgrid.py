"""Cell grids (array-of-structs device storage of whole user cells, the
storage of the generic kernel family): descriptor layout on the CPU; halo
refresh and host <-> device copies on the GPU against a numpy model of the
topology rule (out of bounds along an unwrapped axis -> the edge cell, else
the cell at the coordinate wrapped along the wrapped axes;
src/libgeodecomp/geometry/topologies.h:185-234)."""
import ctypes as C
import itertools

import numpy as np
import pytest

from libgeodecomp_amd import lgdhip


def make(dim, nx, ny=1, nz=1, halo=1, cell_bytes=4):
    g = lgdhip.lgdhip_cell_grid()
    lgdhip.call("lgdhip_cell_grid_make", C.byref(g), dim, nx, ny, nz, halo, cell_bytes)
    return g


def test_cell_grid_layout_without_gpu():
    lgdhip.load()
    g = make(3, 50, 20, 10, halo=2, cell_bytes=28)
    assert (g.px, g.py, g.pz) == (54, 24, 14)
    assert lgdhip.load().lgdhip_cell_grid_bytes(C.byref(g)) == 54 * 24 * 14 * 28
    g = make(2, 121, 151, 99, halo=1, cell_bytes=3)       # nz ignored below 3D
    assert (g.nz, g.px, g.py, g.pz) == (1, 123, 153, 1)
    g = make(1, 777, 5, 5, halo=3, cell_bytes=1)          # ny, nz ignored in 1D
    assert (g.ny, g.nz, g.px, g.py, g.pz) == (1, 1, 783, 1, 1)


@pytest.mark.parametrize("args", [(0, 4, 1, 1, 1, 4), (4, 4, 1, 1, 1, 4), (2, 4, 4, 1, 0, 4), (2, 4, 4, 1, 17, 4),
                                  (2, 0, 4, 1, 1, 4), (2, 4, 4, 1, 1, 0), (3, 4, 4, 0, 1, 8)])
def test_cell_grid_rejects_bad_descriptors(args):
    lib = lgdhip.load()
    g = lgdhip.lgdhip_cell_grid()
    rc = lib.lgdhip_cell_grid_make(C.byref(g), *args)
    assert rc == -1, (args, rc)
    assert lib.lgdhip_last_error()


def halo_model(padded, edge, dim, nx, ny, nz, h, wrap):
    """numpy model: every halo cell of a (pz, py, px, cell_bytes) array."""
    hy, hz = (h if dim >= 2 else 0), (h if dim == 3 else 0)
    out = padded.copy()
    zs, ys, xs = np.meshgrid(np.arange(-hz, nz + hz), np.arange(-hy, ny + hy), np.arange(-h, nx + h), indexing="ij")
    halo = ~((xs >= 0) & (xs < nx) & (ys >= 0) & (ys < ny) & (zs >= 0) & (zs < nz))
    wx = np.mod(xs, nx) if wrap & 1 else xs
    wy = np.mod(ys, ny) if wrap & 2 else ys
    wz = np.mod(zs, nz) if wrap & 4 else zs
    inside = (wx >= 0) & (wx < nx) & (wy >= 0) & (wy < ny) & (wz >= 0) & (wz < nz)
    src = padded[np.clip(wz + hz, 0, None), np.clip(wy + hy, 0, None), np.clip(wx + h, 0, None)]
    val = np.where(inside[..., None], src, edge[None, None, None, :])
    out[halo] = val[halo]
    return out


CASES = [(1, (777, 1, 1), 1, 28), (1, (2, 1, 1), 3, 12), (2, (121, 151, 1), 1, 28), (2, (3, 1, 1), 2, 3),
         (2, (67, 45, 1), 2, 1), (3, (50, 20, 10), 1, 36), (3, (1, 2, 1), 1, 8), (3, (23, 17, 9), 2, 5)]


@pytest.mark.gpu
@pytest.mark.parametrize("dim,shape,h,cb", CASES)
def test_cell_grid_fill_halo_vs_model(hip, dim, shape, h, cb):
    nx, ny, nz = shape
    g = make(dim, nx, ny, nz, halo=h, cell_bytes=cb)
    nbytes = hip.load().lgdhip_cell_grid_bytes(C.byref(g))
    rng = np.random.default_rng(dim * 1000 + cb)
    padded = rng.integers(0, 256, size=(g.pz, g.py, g.px, cb), dtype=np.uint8)
    edge = rng.integers(0, 256, size=cb, dtype=np.uint8)
    buf = hip.DeviceBuffer(nbytes)
    ebuf = hip.DeviceBuffer(cb)
    hip.call("lgdhip_memcpy_h2d", ebuf.ptr, edge.ctypes.data_as(C.c_void_p), cb, None)
    for wrap in range(1 << dim):
        print(f"[case] dim {dim} shape {shape} halo {h} cell_bytes {cb} wrap {wrap}", flush=True)
        hip.call("lgdhip_memcpy_h2d", buf.ptr, padded.ctypes.data_as(C.c_void_p), nbytes, None)
        hip.call("lgdhip_cell_grid_fill_halo", buf.ptr, C.byref(g), ebuf.ptr, wrap, None)
        hip.call("lgdhip_device_synchronize")
        got = np.empty_like(padded)
        hip.call("lgdhip_memcpy_d2h", got.ctypes.data_as(C.c_void_p), buf.ptr, nbytes, None)
        exp = halo_model(padded, edge, dim, nx, ny, nz, h, wrap)
        assert np.array_equal(got, exp), (dim, shape, h, cb, wrap)


@pytest.mark.gpu
@pytest.mark.parametrize("dim,shape,h,cb", CASES)
def test_cell_grid_load_store(hip, dim, shape, h, cb):
    nx, ny, nz = shape
    g = make(dim, nx, ny, nz, halo=h, cell_bytes=cb)
    nbytes = hip.load().lgdhip_cell_grid_bytes(C.byref(g))
    rng = np.random.default_rng(7 + cb)
    host = rng.integers(0, 256, size=(g.nz, g.ny, g.nx, cb), dtype=np.uint8)
    buf = hip.DeviceBuffer(nbytes)
    hip.call("lgdhip_fill_u8", buf.ptr, 0xA5, nbytes, None)
    hip.call("lgdhip_cell_grid_load", buf.ptr, C.byref(g), host.ctypes.data_as(C.c_void_p), None)
    raw = np.empty((g.pz, g.py, g.px, cb), np.uint8)
    hip.call("lgdhip_memcpy_d2h", raw.ctypes.data_as(C.c_void_p), buf.ptr, nbytes, None)
    hy, hz = (h if dim >= 2 else 0), (h if dim == 3 else 0)
    assert np.array_equal(raw[hz:hz + nz, hy:hy + ny, h:h + nx], host)
    mask = np.ones(raw.shape[:3], bool)
    mask[hz:hz + nz, hy:hy + ny, h:h + nx] = False
    assert (raw[mask] == 0xA5).all()          # the halo is untouched
    back = np.zeros_like(host)
    hip.call("lgdhip_cell_grid_store", buf.ptr, C.byref(g), back.ctypes.data_as(C.c_void_p), None)
    assert np.array_equal(back, host)
