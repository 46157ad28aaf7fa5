"""Grid views (lgdhip_grid_view): a sub-box of a grid sharing its memory, the
building block of a region stepper's inner-set / rim split
(src/libgeodecomp/parallelization/nesting/cudastepper.h:255-590 updates the
rim, exchanges it, then updates the inner set).  A single Jacobi sweep over
the rim views and the inner view together equals the full sweep bit for bit;
loads and stores of a view touch only its box; operations that treat the
halo as the edge reject views."""
import ctypes as C

import numpy as np
import pytest

from libgeodecomp_amd import lgdhip


def view(g, x0, y0, z0, nx, ny, nz):
    v = lgdhip.lgdhip_grid()
    lgdhip.call("lgdhip_grid_view", C.byref(g), x0, y0, z0, nx, ny, nz, C.byref(v))
    return v


def test_view_descriptor_without_gpu():
    g = lgdhip.make_grid(3, 37, 29, 23, halo=1, elem_bytes=8)
    v = view(g, 3, 4, 5, 10, 11, 12)
    assert (v.ox, v.oy, v.oz) == (g.ox + 3, g.oy + 4, g.oz + 5)
    assert (v.nx, v.ny, v.nz) == (10, 11, 12) and v.flags == lgdhip.GRID_VIEW
    assert (v.sx, v.sy, v.sz) == (g.sx, g.sy, g.sz)
    lib = lgdhip.load()
    out = lgdhip.lgdhip_grid()
    assert lib.lgdhip_grid_view(C.byref(g), 30, 0, 0, 8, 1, 1, C.byref(out)) == -1       # past nx
    assert lib.lgdhip_grid_view(C.byref(g), 0, 0, 0, 0, 1, 1, C.byref(out)) == -1        # empty


def test_views_rejected_where_the_halo_is_the_edge_without_gpu():
    # every check below returns before any device work
    lib = lgdhip.load()
    fake = [C.c_void_p(0x1000 * (i + 1)) for i in range(3)]
    g = lgdhip.make_grid(3, 37, 29, 23, halo=1, elem_bytes=8)
    v = view(g, 1, 1, 1, 8, 8, 8)
    edge = C.c_double(0.0)
    assert lib.lgdhip_grid_fill_edge(fake[0], C.byref(v), C.byref(edge), None) == -2
    assert lib.lgdhip_grid_fill_halo(fake[0], C.byref(v), C.byref(edge), 1, None) == -2
    assert lib.lgdhip_grid_set_layout(C.byref(v), 1) == -2
    for variant in (4, 5, 6, 7, 8, 9, 10):
        t = lgdhip.tile(variant=variant)
        assert lib.lgdhip_jacobi3d_f64_run(fake[0], fake[1], C.byref(v), 2, C.byref(t), None, None) == -2
    gl = lgdhip.make_grid(3, 16, 16, 16, halo=1, elem_bytes=4, members=19)
    vl = view(gl, 0, 0, 0, 8, 8, 8)
    p = lgdhip.lgdhip_lbm_params(1.7, 0.05, 0.0, 0.0)
    assert lib.lgdhip_lbm_d3q19_f32_run(fake[0], fake[1], fake[2], C.byref(vl), C.byref(p), 1, None, None, None) == -2
    g2 = lgdhip.make_grid(2, 64, 64, 1, halo=1, elem_bytes=1)
    v2 = view(g2, 0, 0, 0, 32, 32, 1)
    assert lib.lgdhip_life2d_u8_run(fake[0], fake[1], C.byref(v2), 1, None, None) == -2


def rim_and_inner(n, w):
    """Boxes (x0, y0, z0, nx, ny, nz) covering an n = (nx, ny, nz) box once:
    the rim of width w (z slabs, then y slabs, then x slabs) and the inner set."""
    nx, ny, nz = n
    boxes = [(0, 0, 0, nx, ny, w), (0, 0, nz - w, nx, ny, w),
             (0, 0, w, nx, w, nz - 2 * w), (0, ny - w, w, nx, w, nz - 2 * w),
             (0, w, w, w, ny - 2 * w, nz - 2 * w), (nx - w, w, w, w, ny - 2 * w, nz - 2 * w)]
    inner = (w, w, w, nx - 2 * w, ny - 2 * w, nz - 2 * w)
    return boxes, inner


def test_rim_inner_partition_covers_once():
    n = (37, 29, 23)
    boxes, inner = rim_and_inner(n, 2)
    cover = np.zeros(n[::-1], np.int32)
    for (x0, y0, z0, bx, by, bz) in boxes + [inner]:
        cover[z0:z0 + bz, y0:y0 + by, x0:x0 + bx] += 1
    assert (cover == 1).all()


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_rim_then_inner_equals_full_sweeps(hip, variant):
    n = (37, 29, 23)
    rng = np.random.default_rng(11)
    host = rng.random(n[::-1])
    g = hip.make_grid(3, *n, halo=1, elem_bytes=8)
    full = [hip.DeviceGrid(g, np.float64) for _ in range(2)]
    split = [hip.DeviceGrid(g, np.float64) for _ in range(2)]
    for d in full + split:
        d.fill_edge(0.5)
        d.load(host)
    tile = hip.tile(variant=variant) if variant else None
    steps = 3
    boxes, inner = rim_and_inner(n, 2)
    print(f"[case] variant {variant}: {steps} steps, rim of width 2, then the inner set", flush=True)
    for t in range(steps):
        cur, nxt = full[t % 2], full[(t + 1) % 2]
        hip.call("lgdhip_jacobi3d_f64_run", cur.ptr, nxt.ptr, C.byref(g), 1,
                 C.byref(tile) if tile else None, None, None)
        cur, nxt = split[t % 2], split[(t + 1) % 2]
        for box in boxes + [inner]:
            v = view(g, *box)
            hip.call("lgdhip_jacobi3d_f64_run", cur.ptr, nxt.ptr, C.byref(v), 1,
                     C.byref(tile) if tile else None, None, None)
    a, b = full[steps % 2].store(), split[steps % 2].store()
    assert np.array_equal(a, b)
    assert not np.array_equal(a, host)


@pytest.mark.gpu
def test_view_load_store(hip):
    n = (37, 29, 23)
    rng = np.random.default_rng(5)
    host = rng.random(n[::-1])
    g = hip.make_grid(3, *n, halo=1, elem_bytes=8)
    d = hip.DeviceGrid(g, np.float64)
    d.fill_edge(0.0)
    d.load(host)
    box = (5, 7, 3, 11, 6, 9)
    v = view(g, *box)
    x0, y0, z0, bx, by, bz = box
    sub = np.empty((bz, by, bx))
    hip.call("lgdhip_grid_store", d.ptr, C.byref(v), sub.ctypes.data_as(C.c_void_p), None)
    assert np.array_equal(sub, host[z0:z0 + bz, y0:y0 + by, x0:x0 + bx])
    new = -rng.random((bz, by, bx))
    hip.call("lgdhip_grid_load", d.ptr, C.byref(v), new.ctypes.data_as(C.c_void_p), None)
    exp = host.copy()
    exp[z0:z0 + bz, y0:y0 + by, x0:x0 + bx] = new
    assert np.array_equal(d.store(), exp)
