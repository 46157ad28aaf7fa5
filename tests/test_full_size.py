"""Parity at the full sizes and step counts of the measured workloads.

* Jacobi3D FP64 128^3 x 100 steps (the measured configuration) and 512^3
  (the HBM-bound shape) bit-exact against the C oracle, through the
  library's default launch plan and the other blocked kernel.
* Game of Life 16384^2 x 1000 generations (the headline run) against the
  SHA-256 and popcount of the oracle's result, written by
  tests/golden/make_life_fixture.py (the oracle needs minutes for this size,
  so the fixture is computed once on the CPU and committed).
* D3Q19 LBM FP32 64^3 x 2000 steps against the FP64 oracle with the
  max-abs errors on rho and u printed and bounded.  The 256^3 x 2000 run is
  opt-in (LGD_FULL_LBM=1): its FP64 oracle takes minutes of CPU.
"""
import ctypes as C
import hashlib
import json
import os
from pathlib import Path

import numpy as np
import pytest

from libgeodecomp_amd import synthetic
from test_gpu_kernels import W, gpu_jacobi, say, shifted

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


# ----------------------------------------------------------------------
# Jacobi3D FP64
# ----------------------------------------------------------------------
@pytest.mark.parametrize("variant", [0, 6, 10])
def test_jacobi_128_cube_100_steps_bit_exact(hip, oracle, variant):
    """The measured configuration: 128^3, U[0,1) from the default seed, zero
    edge, 100 steps; default plan (variant 0) and both blocked kernels."""
    a = synthetic.jacobi_initial(128, 128, 128)
    say("jacobi 128^3 x 100, variant", variant)
    exp = oracle.jacobi3d(a, 100, edge=0.0, threads=8)
    t = None if variant == 0 else hip.tile(variant=variant)
    assert np.array_equal(gpu_jacobi(hip, a, 100, tile=t), exp)


def test_jacobi_512_cube_bit_exact(hip, oracle):
    """The HBM-bound shape, 10 steps (two 4-sweep launches and two single
    sweeps of the default plan)."""
    a = synthetic.jacobi_initial(512, 512, 512, seed=5)
    say("jacobi 512^3 x 10")
    exp = oracle.jacobi3d(a, 10, edge=0.0, threads=8)
    assert np.array_equal(gpu_jacobi(hip, a, 10), exp)


# ----------------------------------------------------------------------
# Game of Life 16384^2 x 1000
# ----------------------------------------------------------------------
def life_fixture():
    return json.loads((GOLDEN / "life_16384x16384.json").read_text())


def digest(grid):
    return hashlib.sha256(np.ascontiguousarray(grid, dtype=np.uint8).tobytes()).hexdigest(), \
        int(np.count_nonzero(grid))


@pytest.fixture(scope="module")
def life_input():
    fx = life_fixture()
    n = fx["grid"][0]
    g0 = synthetic.life_initial(n, n, fx["density"], fx["seed"])
    # the input itself must be the one the fixture was computed from
    assert digest(g0) == (fx["checkpoints"]["0"]["sha256"], fx["checkpoints"]["0"]["popcount"])
    return g0


@pytest.mark.parametrize("tile", [dict(steps_per_launch=31, variant=7), None, dict(steps_per_launch=16, variant=3)])
@pytest.mark.parametrize("generations", [31, 1000])
def test_life_16384_square_against_fixture(hip, life_input, tile, generations):
    """The headline run (the bench's tile: 31 generations per launch,
    stacked-wave kernel; the library default; a one-row-per-thread
    variant) ends in exactly the oracle's state: SHA-256 of all 268 M cells
    and their popcount."""
    fx = life_fixture()["checkpoints"][str(generations)]
    n = life_input.shape[0]
    say("life 16384^2", generations, "generations, tile", tile)
    g = hip.make_grid(2, n, n, 1, halo=1, elem_bytes=1)
    a = hip.DeviceGrid(g, np.uint8)
    a.fill_edge(0)
    a.load(life_input)
    ws_bytes = hip.load().lgdhip_life2d_workspace_bytes(C.byref(g))
    ws = hip.DeviceBuffer(ws_bytes)
    t = hip.tile(**tile) if tile is not None else None
    hip.call("lgdhip_life2d_run", a.ptr, C.byref(g), 0, generations, ws.ptr, ws_bytes,
             C.byref(t) if t is not None else None, None)
    sha, pop = digest(a.store())
    assert pop == fx["popcount"]
    assert sha == fx["sha256"]


# ----------------------------------------------------------------------
# D3Q19 LBM FP32 vs FP64
# ----------------------------------------------------------------------
def lbm_errors(hip, oracle, n, steps, threads, chunk=None):
    """chunk: run the oracle that many steps at a time and report each
    chunk (the oracle's state between steps is f alone, so the result is
    the same as one call)."""
    from test_gpu_kernels import gpu_lbm
    ux, uy, uz = synthetic.lbm_initial_velocity(n, n, n)
    f0 = oracle.lbm_equilibrium(np.ones((n, n, n)), ux, uy, uz)
    g0 = shifted(f0)
    out, flags = gpu_lbm(hip, g0, steps, layout=hip.LAYOUT_Z_SLABS)
    del g0
    chunk = chunk or steps
    ref64 = f0
    for done in range(0, steps, chunk):
        ref64 = oracle.lbm_run(ref64, flags, min(chunk, steps - done), 1.7, (0.05, 0.0, 0.0), threads=threads)
        if chunk < steps:
            say(f"  FP64 oracle {min(done + chunk, steps)}/{steps} steps")
    del f0
    R, JX, JY, JZ = oracle.lbm_moments(ref64)
    del ref64
    f_gpu = out.astype(np.float64) + W.reshape(19, 1, 1, 1)
    del out
    r, jx, jy, jz = oracle.lbm_moments(f_gpu)
    err_rho = float(np.abs(r - R).max())
    err_u = max(float(np.abs(a - b).max()) for a, b in ((jx, JX), (jy, JY), (jz, JZ)))
    umax = max(float(np.abs(b).max()) for b in (JX, JY, JZ))
    return err_rho, err_u, umax


def test_lbm_64_cube_2000_steps_vs_fp64_oracle(hip, oracle):
    """The cavity of the measured configuration at 64^3 for its full 2000
    steps: the FP32 device run (shifted populations, z slabs, the default
    kernel) against the FP64 oracle.  u is the incompressible model's
    velocity, the momentum sum e f.  Bound: 2e-8 per step on both (a few
    FP32 ulps of the deviations |f - w| < 0.05 per step, growing at most
    linearly)."""
    steps = 2000
    say("lbm 64^3 x 2000 vs FP64")
    err_rho, err_u, umax = lbm_errors(hip, oracle, 64, steps, threads=16)
    say(f"lbm 64^3 x {steps}: max-abs rho {err_rho:.3e}, u {err_u:.3e} (max |u| {umax:.3e})")
    assert umax > 1e-2  # the lid has driven a flow
    assert err_rho < 2e-8 * steps and err_u < 2e-8 * steps


@pytest.mark.skipif(os.environ.get("LGD_FULL_LBM") != "1",
                    reason="256^3 x 2000 FP64 oracle takes minutes of CPU: opt in with LGD_FULL_LBM=1")
def test_lbm_256_cube_2000_steps_vs_fp64_oracle(hip, oracle):
    steps = 2000
    say("lbm 256^3 x 2000 vs FP64 (oracle on the host, minutes)")
    err_rho, err_u, umax = lbm_errors(hip, oracle, 256, steps, threads=16, chunk=100)
    say(f"lbm 256^3 x {steps}: max-abs rho {err_rho:.3e}, u {err_u:.3e} (max |u| {umax:.3e})")
    assert err_rho < 2e-8 * steps and err_u < 2e-8 * steps
