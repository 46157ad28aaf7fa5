"""Member layouts of the padded SoA grid (LGDHIP_LAYOUT_MEMBER_PLANES and
LGDHIP_LAYOUT_Z_SLABS): the documented address of element (x, y, z) of
member m, m * MS + (z + oz) * ZS + (y + oy) * sx + (x + ox), checked through
the raw device buffer after lgdhip_grid_load; fill_edge and fill_halo
against a numpy model of the same buffer; load/store round trips."""
import ctypes as C

import numpy as np
import pytest

from libgeodecomp_amd import lgdhip


def strides(g):
    if g.layout == lgdhip.LAYOUT_Z_SLABS:
        return g.sx * g.sy, g.members * g.sx * g.sy
    return g.sx * g.sy * g.sz, g.sx * g.sy


def test_set_layout_without_gpu():
    lib = lgdhip.load()
    g = lgdhip.make_grid(3, 13, 7, 5, halo=2, elem_bytes=4, members=3, layout=lgdhip.LAYOUT_Z_SLABS)
    assert g.layout == lgdhip.LAYOUT_Z_SLABS and (g.nx, g.ny, g.nz) == (13, 7, 5)
    assert lgdhip.grid_bytes(g) == 3 * g.sx * g.sy * g.sz * 4        # same size in both layouts
    assert lib.lgdhip_grid_set_layout(C.byref(g), 7) == -1
    assert "layout" in lib.lgdhip_last_error().decode()


def raw_index(g, m, x, y, z):
    ms, zs = strides(g)
    return m * ms + (z + g.oz) * zs + (y + g.oy) * g.sx + (x + g.ox)


def padded_view(g, raw):
    """(members, sz, sy, sx) view of a raw buffer in either layout."""
    if g.layout == lgdhip.LAYOUT_Z_SLABS:
        return raw.reshape(g.sz, g.members, g.sy, g.sx).transpose(1, 0, 2, 3)
    return raw.reshape(g.members, g.sz, g.sy, g.sx)


@pytest.mark.gpu
@pytest.mark.parametrize("layout", [lgdhip.LAYOUT_MEMBER_PLANES, lgdhip.LAYOUT_Z_SLABS])
@pytest.mark.parametrize("dim,shape,halo,eb,members", [(3, (13, 7, 5), 2, 4, 3), (3, (40, 9, 6), 1, 1, 19),
                                                       (2, (33, 5, 1), 1, 8, 2), (3, (8, 8, 8), 3, 2, 5)])
def test_layout_address_fill_and_round_trip(hip, layout, dim, shape, halo, eb, members):
    nx, ny, nz = shape
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[eb]
    g = hip.make_grid(dim, nx, ny, nz, halo=halo, elem_bytes=eb, members=members, layout=layout)
    print(f"[case] layout {layout} dim {dim} shape {shape} halo {halo} eb {eb} members {members}", flush=True)
    d = hip.DeviceGrid(g, dt)
    n = hip.grid_bytes(g) // eb
    rng = np.random.default_rng(members * 100 + eb)
    edge = rng.integers(1, 200, size=members).astype(dt)
    d.fill_edge(edge)
    host = rng.integers(0, 2 ** (8 * eb - 1), size=d.logical_shape(), dtype=np.uint64).astype(dt)
    d.load(host)
    raw = np.empty(n, dt)
    hip.call("lgdhip_memcpy_d2h", raw.ctypes.data_as(C.c_void_p), d.ptr, n * eb, None)
    hm = host.reshape((members, nz, ny, nx))
    # the documented address formula, probed at the corners and a few inner points
    pts = [(0, 0, 0), (nx - 1, ny - 1, nz - 1), (nx // 2, ny // 3, nz // 2), (1 % nx, 0, nz - 1)]
    for m in range(members):
        for (x, y, z) in pts:
            assert raw[raw_index(g, m, x, y, z)] == hm[m, z, y, x], (m, x, y, z)
    # whole buffer: interior = host, every other element = the member's edge
    pv = padded_view(g, raw)
    oz = g.oz
    exp = np.broadcast_to(edge.reshape(members, 1, 1, 1), pv.shape).copy()
    exp[:, oz:oz + nz, g.oy:g.oy + ny, g.ox:g.ox + nx] = hm
    assert np.array_equal(pv, exp)
    assert np.array_equal(d.store(), host)
    # periodic halo refresh: wrapped axes copy the opposite interior layers
    for wrap in range(1 << dim):
        hip.call("lgdhip_grid_fill_edge", d.ptr, C.byref(g), edge.ctypes.data_as(C.c_void_p), None)
        d.load(host)
        hip.call("lgdhip_grid_fill_halo", d.ptr, C.byref(g), edge.ctypes.data_as(C.c_void_p), wrap, None)
        hip.call("lgdhip_device_synchronize")
        hip.call("lgdhip_memcpy_d2h", raw.ctypes.data_as(C.c_void_p), d.ptr, n * eb, None)
        pv = padded_view(g, raw)
        h = halo
        zr = range(-h, nz + h) if dim == 3 else range(0, 1)
        for m in range(members):
            for z in zr:
                for y in range(-h, ny + h):
                    xs = np.arange(-h, nx + h)
                    wx = np.mod(xs, nx) if wrap & 1 else xs
                    wy = y % ny if wrap & 2 else y
                    wz = z % nz if (wrap & 4) else z
                    inside = (wx >= 0) & (wx < nx) & (0 <= wy < ny) & (0 <= wz < nz)
                    row = np.where(inside, hm[m, min(max(wz, 0), nz - 1), min(max(wy, 0), ny - 1),
                                              np.clip(wx, 0, nx - 1)], edge[m])
                    got = pv[m, z + oz, y + g.oy, g.ox - h:g.ox + nx + h]
                    assert np.array_equal(got, row), (wrap, m, z, y)
