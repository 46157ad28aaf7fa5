"""GPU parity of the liblgdhip kernel families against the CPU oracle.

All calls go through the C-ABI (ctypes).  Integer work (Life) and the
Jacobi FP64 sweep are compared bit-exactly; the FP32 LBM is compared with
the FP32 oracle to a few ulps and with the FP64 oracle to stated max-abs
errors on rho and j.
"""
import numpy as np
import pytest

from libgeodecomp_amd import synthetic
from test_oracle import encode, jacobi_kat_expected, jacobi_kat_input, place

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


# ----------------------------------------------------------------------
# Jacobi3D FP64
# ----------------------------------------------------------------------
def gpu_jacobi(hip, host, steps, edge=0.0, tile=None):
    nz, ny, nx = host.shape
    g = hip.make_grid(3, nx, ny, nz, halo=1, elem_bytes=8)
    a, b = hip.DeviceGrid(g, np.float64), hip.DeviceGrid(g, np.float64)
    for d in (a, b):
        d.fill_edge(edge)
    a.load(host)
    import ctypes as C
    in_b = C.c_int(-1)
    t = C.byref(tile) if tile is not None else None
    hip.call("lgdhip_jacobi3d_f64_run", a.ptr, b.ptr, C.byref(g), steps, t, C.byref(in_b), None)
    hip.call("lgdhip_device_synchronize")
    # one launch swaps the buffers once: K sweeps per launch for the
    # temporal-blocking variants (and the default, which blocks 4 sweeps),
    # so the result's buffer follows the launches
    if tile is not None and (1 <= tile.variant <= 3 or tile.variant == 13):
        assert in_b.value == steps % 2
    assert in_b.value in (0, 1)
    return (b if in_b.value else a).store()


def test_jacobi_known_answer(hip):
    origin, dims = (10, 20, 30), (200, 100, 50)
    say("jacobi KAT", dims)
    out = gpu_jacobi(hip, jacobi_kat_input(origin, dims), 1, edge=encode(-1, -1, -1))
    assert np.array_equal(out, jacobi_kat_expected(origin, dims))


@pytest.mark.parametrize("shape,steps", [((128, 128, 128), 7), ((67, 33, 45), 5), ((1, 1, 1), 3), ((200, 3, 2), 4)])
def test_jacobi_bit_exact_vs_oracle(hip, oracle, shape, steps):
    nx, ny, nz = shape
    say("jacobi", shape, steps)
    a = synthetic.jacobi_initial(nx, ny, nz)
    exp = oracle.jacobi3d(a, steps, edge=0.0, threads=4)
    assert np.array_equal(gpu_jacobi(hip, a, steps), exp)


def test_jacobi_tiles_agree(hip, oracle):
    a = synthetic.jacobi_initial(96, 70, 40, seed=9)
    exp = oracle.jacobi3d(a, 6, edge=0.5, threads=4)
    for variant in (1, 2, 3, 13):
        for by in (1, 2, 4, 8):
            for bz in (1, 3, 9, 64):
                say("jacobi tile", variant, by, bz)
                t = hip.tile(by=by, bz=bz, variant=variant)
                assert np.array_equal(gpu_jacobi(hip, a, 6, edge=0.5, tile=t), exp)


@pytest.mark.parametrize("shape", [(97, 70, 40), (1, 5, 3), (129, 3, 17), (128, 9, 6)])
def test_jacobi_pair_march_bit_exact(hip, oracle, shape):
    """Variant 13 (two cells per thread, 16-byte loads, the x neighbours as
    unaligned pairs): odd widths leave a last pair whose second cell is not
    stored; a one-cell row; widths at and past one 128-cell segment."""
    nx, ny, nz = shape
    a = synthetic.jacobi_initial(nx, ny, nz, seed=nx + ny)
    exp = oracle.jacobi3d(a, 5, edge=0.25, threads=4)
    for bz in (2, 64):
        say("jacobi pair", shape, bz)
        t = hip.tile(by=4, bz=bz, variant=13)
        assert np.array_equal(gpu_jacobi(hip, a, 5, edge=0.25, tile=t), exp)


@pytest.mark.parametrize("shape,steps", [((96, 70, 40), 7), ((130, 17, 9), 9), ((1, 1, 1), 5), ((61, 1, 30), 4),
                                         ((64, 64, 64), 8)])
def test_jacobi_temporal_blocking_bit_exact(hip, oracle, shape, steps):
    """K sweeps per launch (variants 4-8: K = 2, 3, 4 with 16-row tiles, K = 2,
    3 with 8-row tiles; 9, 10: K = 4 with two rows per thread, 16- and
    32-row tiles), step counts that leave a remainder of single
    sweeps, ragged shapes, z chunks down to one plane and chunks longer than
    the kernel's 6-step unrolled march, non-zero edge; 11, 12: the same
    pipelines for one sweep per launch."""
    nx, ny, nz = shape
    a = synthetic.jacobi_initial(nx, ny, nz, seed=nx + ny + nz)
    exp = oracle.jacobi3d(a, steps, edge=0.5, threads=4)
    for variant in (4, 5, 6, 7, 8, 9, 10, 11, 12):
        for bz in (0, 1, 5, 13):
            say("jacobi tb", shape, steps, variant, bz)
            t = hip.tile(bz=bz, variant=variant)
            assert np.array_equal(gpu_jacobi(hip, a, steps, edge=0.5, tile=t), exp)


def gpu_jacobi_periodic(hip, host, steps, wrap, edge=0.0, tile=None, stale=(77.0, 99.0)):
    """lgdhip_jacobi3d_f64_run_periodic; the buffers' halos start out holding
    other values than `edge` (the entry point rewrites both)."""
    import ctypes as C
    nz, ny, nx = host.shape
    g = hip.make_grid(3, nx, ny, nz, halo=1, elem_bytes=8)
    a, b = hip.DeviceGrid(g, np.float64), hip.DeviceGrid(g, np.float64)
    a.fill_edge(stale[0])
    b.fill_edge(stale[1])
    a.load(host)
    in_b = C.c_int(-1)
    t = C.byref(tile) if tile is not None else None
    hip.call("lgdhip_jacobi3d_f64_run_periodic", a.ptr, b.ptr, C.byref(g), steps, wrap, edge, t, C.byref(in_b), None)
    hip.call("lgdhip_device_synchronize")
    assert in_b.value in (0, 1)
    return (b if in_b.value else a).store()


@pytest.mark.parametrize("shape,steps", [((96, 70, 40), 9), ((37, 22, 15), 12), ((10, 5, 9), 8), ((33, 20, 5), 5),
                                         ((1, 1, 1), 4), ((130, 33, 8), 4)])
def test_jacobi_periodic_bit_exact(hip, oracle, shape, steps):
    """Torus and selectively wrapped axes (every wrap mask) through the
    temporal-blocking kernels with wrapped tile coordinates: the default
    (K = 4, one or two rows per thread), K = 2 and 3, a single-sweep variant,
    step counts with a remainder of halo-refresh single sweeps, boxes
    narrower than a tile (10 x 5), a wrapped z axis too short for blocking
    (nz = 5: single sweeps throughout) and exactly long enough (nz = 8)."""
    nx, ny, nz = shape
    a = synthetic.jacobi_initial(nx, ny, nz, seed=nx * ny + nz)
    for wrap in range(8):
        exp = oracle.jacobi3d(a, steps, edge=0.5, threads=4, wrap=wrap)
        for variant, bz in ((0, 0), (10, 0), (6, 5), (4, 0), (8, 3), (2, 0), (11, 0), (12, 7)):
            say("jacobi periodic", shape, steps, "wrap", wrap, "variant", variant, "bz", bz)
            t = hip.tile(bz=bz, variant=variant) if variant else None
            got = gpu_jacobi_periodic(hip, a, steps, wrap, edge=0.5, tile=t)
            assert np.array_equal(got, exp), (wrap, variant, bz)


def test_jacobi_zrange_launch_writes_only_its_planes(hip, oracle):
    """lgdhip_jacobi3d_f64_run_zrange: one blocked launch writes the planes
    [z0, z1) of dst, equal to K oracle sweeps there, and leaves every other
    plane of dst as it was (the rim / inner split of the slab stepper)."""
    import ctypes as C
    nx, ny, nz = 70, 45, 40
    a = synthetic.jacobi_initial(nx, ny, nz, seed=5)
    sentinel = np.full((nz, ny, nx), -7.25)
    g = hip.make_grid(3, nx, ny, nz, halo=1, elem_bytes=8)
    src, dst = hip.DeviceGrid(g, np.float64), hip.DeviceGrid(g, np.float64)
    for d in (src, dst):
        d.fill_edge(0.5)
    src.load(a)
    for variant, k in ((0, 4), (4, 2), (5, 3), (11, 1), (10, 4), (12, 1)):
        exp = oracle.jacobi3d(a, k, edge=0.5, threads=4)
        for z0, z1 in ((0, 4), (4, 36), (36, 40), (13, 14), (0, 40)):
            say("jacobi zrange", variant, z0, z1)
            dst.load(sentinel)
            sweeps = C.c_int(0)
            hip.call("lgdhip_jacobi3d_f64_run_zrange", src.ptr, dst.ptr, C.byref(g), z0, z1,
                     C.byref(hip.tile(variant=variant)), C.byref(sweeps), None)
            out = dst.store()
            assert sweeps.value == k
            assert np.array_equal(out[z0:z1], exp[z0:z1])
            assert np.array_equal(out[:z0], sentinel[:z0]) and np.array_equal(out[z1:], sentinel[z1:])


def test_jacobi_periodic_zero_steps(hip):
    a = synthetic.jacobi_initial(12, 10, 9, seed=3)
    assert np.array_equal(gpu_jacobi_periodic(hip, a, 0, 7), a)


# ----------------------------------------------------------------------
# Game of Life
# ----------------------------------------------------------------------
def gpu_life_u8(hip, host, steps, edge=0):
    import ctypes as C
    ny, nx = host.shape
    g = hip.make_grid(2, nx, ny, 1, halo=1, elem_bytes=1)
    a, b = hip.DeviceGrid(g, np.uint8), hip.DeviceGrid(g, np.uint8)
    for d in (a, b):
        d.fill_edge(edge)
    a.load(host)
    in_b = C.c_int(-1)
    hip.call("lgdhip_life2d_u8_run", a.ptr, b.ptr, C.byref(g), steps, C.byref(in_b), None)
    return (b if in_b.value else a).store()


def gpu_life_packed(hip, host, steps, edge=0, tile=None):
    import ctypes as C
    ny, nx = host.shape
    g = hip.make_grid(2, nx, ny, 1, halo=1, elem_bytes=1)
    a = hip.DeviceGrid(g, np.uint8)
    a.fill_edge(edge)
    a.load(host)
    ws_bytes = hip.load().lgdhip_life2d_workspace_bytes(C.byref(g))
    ws = hip.DeviceBuffer(ws_bytes)
    t = C.byref(tile) if tile is not None else None
    hip.call("lgdhip_life2d_run", a.ptr, C.byref(g), edge, steps, ws.ptr, ws_bytes, t, None)
    return a.store()


LIFE_SHAPES = [(64, 64), (61, 37), (1, 1), (33, 5), (4100, 130), (97, 1000)]


@pytest.mark.parametrize("shape", LIFE_SHAPES)
@pytest.mark.parametrize("edge", [0, 1])
def test_life_u8_swar_bit_exact(hip, oracle, shape, edge):
    nx, ny = shape
    say("life u8", shape, edge)
    g0 = synthetic.life_initial(nx, ny, seed=nx * 7 + ny)
    assert np.array_equal(gpu_life_u8(hip, g0, 13, edge), oracle.life2d(g0, 13, edge=edge, threads=4))


@pytest.mark.parametrize("shape", LIFE_SHAPES)
@pytest.mark.parametrize("edge", [0, 1])
def test_life_packed_bit_exact(hip, oracle, shape, edge):
    nx, ny = shape
    say("life packed", shape, edge)
    g0 = synthetic.life_initial(nx, ny, seed=nx * 3 + ny)
    exp = oracle.life2d(g0, 45, edge=edge, threads=4)
    assert np.array_equal(gpu_life_packed(hip, g0, 45, edge), exp)


@pytest.mark.parametrize("variant", list(range(0, 9)))
@pytest.mark.parametrize("per_launch", [1, 7, 16, 31])
@pytest.mark.parametrize("shape", [(300, 211), (130, 1500)])
def test_life_packed_variants(hip, oracle, variant, per_launch, shape):
    nx, ny = shape
    say("life variant", variant, per_launch, shape)
    g0 = synthetic.life_initial(nx, ny, seed=variant * 100 + per_launch)
    exp = oracle.life2d(g0, 40, threads=4)
    out = gpu_life_packed(hip, g0, 40, tile=hip.tile(steps_per_launch=per_launch, variant=variant))
    assert np.array_equal(out, exp)


def gpu_life_packed_periodic(hip, host, steps, wrap, edge=0, tile=None):
    import ctypes as C
    ny, nx = host.shape
    g = hip.make_grid(2, nx, ny, 1, halo=1, elem_bytes=1)
    a = hip.DeviceGrid(g, np.uint8)
    a.fill_edge(edge)
    a.load(host)
    ws_bytes = hip.load().lgdhip_life2d_workspace_bytes(C.byref(g))
    ws = hip.DeviceBuffer(ws_bytes)
    pa = ws.ptr.value
    pb = pa + ws_bytes // 2
    hip.call("lgdhip_life2d_pack", a.ptr, C.byref(g), C.c_void_p(pa), None)
    in_b = C.c_int(-1)
    t = C.byref(tile) if tile is not None else None
    hip.call("lgdhip_life2d_packed_run_periodic", C.c_void_p(pa), C.c_void_p(pb), C.byref(g), edge, steps, wrap, t,
             C.byref(in_b), None)
    hip.call("lgdhip_life2d_unpack", C.c_void_p(pb if in_b.value else pa), C.byref(g), a.ptr, None)
    return a.store()


@pytest.mark.parametrize("shape", [(64, 64), (96, 45), (32, 7), (4096, 130), (320, 1500), (61, 37), (300, 211)])
def test_life_packed_periodic_bit_exact(hip, oracle, shape):
    """Torus and cylinder topologies through the bit-packed blocked kernels:
    rows wrap for any height (7 rows: a band wraps several times), word
    columns wrap when nx % 32 == 0 (32 cells: one word, its own neighbour),
    both edge states on the unwrapped axis, the default and the other
    kernel variants, launch depths with a remainder."""
    nx, ny = shape
    g0 = synthetic.life_initial(nx, ny, seed=nx + 7 * ny)
    for wrap in ((1, 2, 3) if nx % 32 == 0 else (2,)):
        for edge in (0, 1):
            exp = oracle.life2d(g0, 45, edge=edge, threads=4, wrap=wrap)
            for variant, per_launch in ((0, 0), (7, 16), (1, 31), (3, 7), (5, 1), (8, 31), (2, 13)):
                say("life periodic", shape, "wrap", wrap, "edge", edge, "variant", variant, "per launch", per_launch)
                t = hip.tile(steps_per_launch=per_launch, variant=variant) if variant else None
                out = gpu_life_packed_periodic(hip, g0, 45, wrap, edge=edge, tile=t)
                assert np.array_equal(out, exp), (wrap, edge, variant, per_launch)


def test_life_glider_on_gpu(hip):
    glider = [(11, 10), (12, 11), (10, 12), (11, 12), (12, 12)]
    g = place(glider, (300, 70))
    say("life glider")
    exp = place([(x + 40, y + 40) for x, y in glider], (300, 70))
    assert np.array_equal(gpu_life_packed(hip, g, 160), exp)
    assert np.array_equal(gpu_life_u8(hip, g, 160), exp)


def test_life_packed_rejects_bad_tile(hip):
    with pytest.raises(hip.LgdHipError):
        gpu_life_packed(hip, np.zeros((8, 8), np.uint8), 3, tile=hip.tile(steps_per_launch=33))


# ----------------------------------------------------------------------
# D3Q19 LBM
# ----------------------------------------------------------------------
def gpu_lbm(hip, f_host, steps, omega=1.7, lid=(0.05, 0.0, 0.0), tile=None, layout=0):
    import ctypes as C
    _, nz, ny, nx = f_host.shape
    g = hip.make_grid(3, nx, ny, nz, halo=1, elem_bytes=4, members=19, layout=layout)
    fg = hip.make_grid(3, nx, ny, nz, halo=1, elem_bytes=1, members=1)
    a, b = hip.DeviceGrid(g, np.float32), hip.DeviceGrid(g, np.float32)
    fl = hip.DeviceGrid(fg, np.uint8)
    fl.fill_edge(0)
    hip.call("lgdhip_lbm_cavity_flags", fl.ptr, C.byref(fg), None)
    for d in (a, b):
        d.fill_edge(np.zeros(19, np.float32))
    a.load(f_host.astype(np.float32))
    p = hip.lgdhip_lbm_params(omega, *lid)
    in_b = C.c_int(-1)
    hip.call("lgdhip_lbm_d3q19_f32_run", a.ptr, b.ptr, fl.ptr, C.byref(g), C.byref(p), steps,
             C.byref(tile) if tile is not None else None, C.byref(in_b), None)
    return (b if in_b.value else a).store(), fl.store()


W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def shifted(f):
    """Device storage: g = f - w per population (see include/lgdhip.h)."""
    return (f - W.reshape((19,) + (1,) * (f.ndim - 1))).astype(np.float32)


@pytest.mark.parametrize("shape,steps", [((32, 32, 32), 30), ((17, 9, 23), 25), ((24, 20, 16), 300)])
def test_lbm_vs_oracle(hip, oracle, shape, steps):
    nx, ny, nz = shape
    say("lbm", shape, steps)
    ux, uy, uz = synthetic.lbm_initial_velocity(nx, ny, nz)
    f0 = oracle.lbm_equilibrium(np.ones((nz, ny, nx)), ux, uy, uz)
    g0 = shifted(f0)
    out, flags = gpu_lbm(hip, g0, steps)
    assert np.array_equal(flags, oracle.lbm_cavity_flags(nx, ny, nz))
    # the cell's update maps shifted populations to shifted populations, so
    # the FP32 oracle runs the same arithmetic on g
    ref32 = oracle.lbm_run(g0, flags, steps, 1.7, (0.05, 0.0, 0.0), threads=4)
    ref64 = oracle.lbm_run(f0, flags, steps, 1.7, (0.05, 0.0, 0.0), threads=4)
    # same FP32 formula, only FMA contraction differs
    assert np.abs(out - ref32).max() < 2e-7

    R, JX, JY, JZ = oracle.lbm_moments(ref64)
    f_gpu = out.astype(np.float64) + W.reshape(19, 1, 1, 1)
    r, jx, jy, jz = oracle.lbm_moments(f_gpu)
    err_rho = np.abs(r - R).max()
    err_j = max(np.abs(a - b).max() for a, b in ((jx, JX), (jy, JY), (jz, JZ)))
    err_f = np.abs(f_gpu - ref64).max()
    say(f"lbm FP32 (shifted) vs FP64 oracle after {steps} steps: max-abs rho {err_rho:.3e}  j {err_j:.3e}  "
        f"f {err_f:.3e}")
    # tolerance: shifted FP32 keeps rho and j to a few ulps of the
    # deviations (|g| < 0.05), growing at most linearly in the step count
    assert err_rho < 2e-8 * steps and err_j < 2e-8 * steps and err_f < 2e-8 * steps
    # mass (sum of g over the cavity) is conserved by bounce-back and the lid
    assert abs(out.astype(np.float64).sum() - g0.astype(np.float64).sum()) < 1e-6 * out[0].size


def test_lbm_tile_variants_identical(hip, oracle):
    """Work-group shapes (64 x 1/2/4/8, 256 x 1), nontemporal loads, four
    cells per thread (variant 2), two steps per pass (variant 3: 7 steps =
    3 fused pairs + 1 single step, z chunks down to 4 planes) and the three
    tile orders change the schedule, not the per-cell arithmetic: all results equal the default
    launch's bit for bit, on a grid ragged in every direction (nx = 300: the
    last 256-cell segment of a row is partly used)."""
    nx, ny, nz = 300, 37, 11
    ux, uy, uz = synthetic.lbm_initial_velocity(nx, ny, nz)
    g0 = shifted(oracle.lbm_equilibrium(np.ones((nz, ny, nx)), ux, uy, uz))
    ref, _ = gpu_lbm(hip, g0, 7)
    for kw in (dict(by=1), dict(by=2), dict(by=8), dict(bx=256), dict(variant=1), dict(flags=hip.TILE_XCD_SWIZZLE),
               dict(flags=hip.TILE_XCD_ROWS), dict(bx=256, flags=hip.TILE_XCD_SWIZZLE), dict(variant=2),
               dict(variant=2, by=2), dict(variant=2, by=4, flags=hip.TILE_XCD_SWIZZLE),
               dict(variant=2, flags=hip.TILE_XCD_ROWS), dict(variant=3), dict(variant=3, by=2),
               dict(variant=3, bz=4), dict(variant=3, by=2, bz=5)):
        say("lbm tile", kw)
        out, _ = gpu_lbm(hip, g0, 7, tile=hip.tile(**kw))
        assert np.array_equal(out, ref), kw


def test_lbm_z_slab_layout_identical(hip, oracle):
    """The z-slab member layout changes addresses only: every launch shape
    gives the member-plane layout's result bit for bit."""
    nx, ny, nz = 300, 37, 11
    ux, uy, uz = synthetic.lbm_initial_velocity(nx, ny, nz)
    g0 = shifted(oracle.lbm_equilibrium(np.ones((nz, ny, nx)), ux, uy, uz))
    ref, _ = gpu_lbm(hip, g0, 7)
    for kw in (dict(), dict(by=2), dict(bx=256), dict(variant=2), dict(flags=hip.TILE_XCD_SWIZZLE), dict(variant=3),
               dict(variant=3, by=2, bz=3)):
        say("lbm z slabs", kw)
        out, _ = gpu_lbm(hip, g0, 7, tile=hip.tile(**kw) if kw else None, layout=hip.LAYOUT_Z_SLABS)
        assert np.array_equal(out, ref), kw


def test_lbm_rest_state(hip, oracle):
    n = (16, 8, 8)
    f0 = oracle.lbm_equilibrium(np.ones(n[::-1]), *(np.zeros(n[::-1]),) * 3)
    out, fl = gpu_lbm(hip, shifted(f0), 20, lid=(0.0, 0.0, 0.0))
    # in shifted storage the rest state is exactly zero and stays zero
    assert np.abs(out).max() < 1e-12
