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
    assert in_b.value == steps % 2
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
    for by in (1, 2, 4, 8):
        for bz in (1, 3, 64):
            say("jacobi tile", by, bz)
            assert np.array_equal(gpu_jacobi(hip, a, 6, edge=0.5, tile=hip.tile(by=by, bz=bz)), exp)


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


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("per_launch", [1, 7, 16, 31])
def test_life_packed_variants(hip, oracle, variant, per_launch):
    say("life variant", variant, per_launch)
    g0 = synthetic.life_initial(300, 211, seed=variant * 100 + per_launch)
    exp = oracle.life2d(g0, 40, threads=4)
    out = gpu_life_packed(hip, g0, 40, tile=hip.tile(steps_per_launch=per_launch, variant=variant))
    assert np.array_equal(out, exp)


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
def gpu_lbm(hip, f_host, steps, omega=1.7, lid=(0.05, 0.0, 0.0)):
    import ctypes as C
    _, nz, ny, nx = f_host.shape
    g = hip.make_grid(3, nx, ny, nz, halo=1, elem_bytes=4, members=19)
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
    hip.call("lgdhip_lbm_d3q19_f32_run", a.ptr, b.ptr, fl.ptr, C.byref(g), C.byref(p), steps, None,
             C.byref(in_b), None)
    return (b if in_b.value else a).store(), fl.store()


@pytest.mark.parametrize("shape,steps", [((32, 32, 32), 30), ((17, 9, 23), 25)])
def test_lbm_vs_oracle(hip, oracle, shape, steps):
    nx, ny, nz = shape
    say("lbm", shape, steps)
    ux, uy, uz = synthetic.lbm_initial_velocity(nx, ny, nz)
    f0 = oracle.lbm_equilibrium(np.ones((nz, ny, nx)), ux, uy, uz)
    out, flags = gpu_lbm(hip, f0, steps)
    assert np.array_equal(flags, oracle.lbm_cavity_flags(nx, ny, nz))
    ref32 = oracle.lbm_run(f0.astype(np.float32), flags, steps, 1.7, (0.05, 0.0, 0.0), threads=4)
    ref64 = oracle.lbm_run(f0, flags, steps, 1.7, (0.05, 0.0, 0.0), threads=4)
    # same FP32 formula, only contraction/ordering differs: a few ulps of 1/18
    assert np.abs(out - ref32).max() < 1e-6

    def moment_errors(f):
        r, jx, jy, jz = oracle.lbm_moments(f.astype(np.float64))
        R, JX, JY, JZ = oracle.lbm_moments(ref64)
        return np.abs(r - R).max(), max(np.abs(a - b).max() for a, b in ((jx, JX), (jy, JY), (jz, JZ)))

    err_rho, err_j = moment_errors(out)
    o_rho, o_j = moment_errors(ref32)
    say(f"lbm max-abs vs FP64 oracle after {steps} steps: rho {err_rho:.3e} j {err_j:.3e} "
        f"(FP32 oracle: rho {o_rho:.3e} j {o_j:.3e})")
    # FP32 storage: rho = sum of 19 values near 1/19..1/3 rounds at 2^-24
    # per add, so the error grows at most ~19*2^-24 per step; the kernel
    # may not be worse than the FP32 restatement of the same cell by more
    # than a rounding-order margin
    assert err_rho <= 19 * 2.0 ** -24 * steps and err_j <= 19 * 2.0 ** -24 * steps
    assert err_rho <= 1.5 * o_rho + 1e-6 and err_j <= 1.5 * o_j + 1e-7
    # mass is conserved by bounce-back and the lid term
    assert abs(out.astype(np.float64).sum() - f0.sum()) < 1e-4


def test_lbm_rest_state(hip, oracle):
    n = (16, 8, 8)
    f0 = oracle.lbm_equilibrium(np.ones(n[::-1]), *(np.zeros(n[::-1]),) * 3)
    out, fl = gpu_lbm(hip, f0, 20, lid=(0.0, 0.0, 0.0))
    ref32 = oracle.lbm_run(f0.astype(np.float32), fl, 20, 1.7, (0.0, 0.0, 0.0))
    # the rest state is a fixed point up to FP32 rounding of rho = sum f
    assert np.abs(out - f0).max() < 20 * 19 * 2.0 ** -24
    assert np.abs(out - ref32).max() < 1e-7
