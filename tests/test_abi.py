"""The C-ABI library loads and exports every symbol include/lgdhip.h declares
(CPU only: no compute call is made)."""
import ctypes as C
import re
from pathlib import Path

from libgeodecomp_amd import lgdhip

REPO = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (REPO / "include" / "lgdhip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lgdhip_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "lgdhip_jacobi3d_f64_run" in names and "lgdhip_life2d_run" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(lgdhip.library_path()))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) == set(lgdhip.SIGNATURES)


def test_grid_descriptor_without_gpu():
    # pure host logic of the C-ABI: descriptor layout and error reporting
    g = lgdhip.make_grid(3, 130, 7, 5, halo=1, elem_bytes=8)
    assert (g.ox * 8) % 128 == 0 and (g.sx * 8) % 128 == 0
    assert g.sx >= g.ox + g.nx + 1 and g.sy == 9 and g.sz == 7
    assert lgdhip.grid_bytes(g) == g.sx * g.sy * g.sz * 8
    g2 = lgdhip.make_grid(2, 33, 5, 1, halo=1, elem_bytes=1)
    assert g2.sz == 1 and g2.oz == 0 and g2.sx >= g2.ox + 64 + 1
    bad = lgdhip.lgdhip_grid()
    try:
        lgdhip.call("lgdhip_grid_make", C.byref(bad), 4, 1, 1, 1, 1, 8, 1)
    except lgdhip.LgdHipError as e:
        assert e.code == -1 and "dim" in str(e)
    else:
        raise AssertionError("invalid dim accepted")
    lib = lgdhip.load()
    assert lib.lgdhip_error_string(-1) == b"invalid argument"
    assert lib.lgdhip_life2d_packed_words(C.byref(g2)) == 2 * 5


def test_struct_layouts_match_header():
    assert C.sizeof(lgdhip.lgdhip_grid) == 6 * 4 + 9 * 8
    assert lgdhip.lgdhip_grid.nx.offset == 24
    assert C.sizeof(lgdhip.lgdhip_cell_grid) == 2 * 4 + 7 * 8
    assert C.sizeof(lgdhip.lgdhip_tile) == 6 * 4
    assert C.sizeof(lgdhip.lgdhip_lbm_params) == 4 * 4


def test_lbm_x4_variant_rejects_ragged_rows_without_gpu():
    # four cells per thread needs nx % 4 == 0; the check comes before any
    # device work, so it runs on the host
    lib = lgdhip.load()
    g = lgdhip.make_grid(3, 30, 8, 8, halo=1, elem_bytes=4, members=19)
    p = lgdhip.lgdhip_lbm_params(1.7, 0.05, 0.0, 0.0)
    t = lgdhip.tile(variant=2)
    fake = C.c_void_p(0x1000), C.c_void_p(0x2000), C.c_void_p(0x3000)
    rc = lib.lgdhip_lbm_d3q19_f32_run(fake[0], fake[1], fake[2], C.byref(g), C.byref(p), 1, C.byref(t), None, None)
    assert rc == -2 and "nx % 4" in lib.lgdhip_last_error().decode()


def test_jacobi_periodic_rejects_bad_arguments_without_gpu():
    # argument checks come before any device work
    lib = lgdhip.load()
    g = lgdhip.make_grid(3, 12, 10, 9, halo=1, elem_bytes=8)
    fake = C.c_void_p(0x1000), C.c_void_p(0x2000)
    in_b = C.c_int(-1)
    rc = lib.lgdhip_jacobi3d_f64_run_periodic(fake[0], fake[1], C.byref(g), 4, 8, 0.0, None, C.byref(in_b), None)
    assert rc == -1 and "wrap_mask" in lib.lgdhip_last_error().decode()
    rc = lib.lgdhip_jacobi3d_f64_run_periodic(fake[0], fake[1], C.byref(g), -1, 7, 0.0, None, C.byref(in_b), None)
    assert rc == -1 and "steps" in lib.lgdhip_last_error().decode()
    v = lgdhip.lgdhip_grid()
    assert lib.lgdhip_grid_view(C.byref(g), 1, 1, 1, 4, 4, 4, C.byref(v)) == 0
    rc = lib.lgdhip_jacobi3d_f64_run_periodic(fake[0], fake[1], C.byref(v), 4, 7, 0.0, None, C.byref(in_b), None)
    assert rc == -2 and "view" in lib.lgdhip_last_error().decode()


def test_life_periodic_rejects_ragged_wrapped_x_without_gpu():
    # a wrapped x axis needs whole packed words; checked before device work
    lib = lgdhip.load()
    g = lgdhip.make_grid(2, 100, 40, 1, halo=1, elem_bytes=1)
    fake = C.c_void_p(0x1000), C.c_void_p(0x2000)
    in_b = C.c_int(-1)
    rc = lib.lgdhip_life2d_packed_run_periodic(fake[0], fake[1], C.byref(g), 0, 4, 1, None, C.byref(in_b), None)
    assert rc == -2 and "nx % 32" in lib.lgdhip_last_error().decode()
    rc = lib.lgdhip_life2d_packed_run_periodic(fake[0], fake[1], C.byref(g), 0, 4, 4, None, C.byref(in_b), None)
    assert rc == -1 and "wrap_mask" in lib.lgdhip_last_error().decode()


def test_conv_forward_workspace_without_gpu():
    """The split partials' workspace (host-only query and check): the small
    levels of the U-Net ask for one, the 64^3 and 128^3 levels (LDS-tile
    and z-marching kernels) do not; a workspace smaller than asked for is
    refused before any launch."""
    lib = lgdhip.load()
    q = lib.lgdhip_conv3d_forward_workspace_bytes
    assert q(32, 192, 3, 64) > 0 and q(16, 128, 3, 128) > 0 and q(8, 192, 3, 64) > 0
    assert q(128, 16, 3, 16) == 0 and q(64, 32, 3, 32) == 0 and q(16, 3, 3, 16) == 0 and q(16, 16, 1, 1) == 0
    assert q(16, 64, 3, 16) == 0  # 16 output channels: not the GEMM kernel
    fake = C.c_void_p(0x1000)
    rc = lib.lgdhip_conv3d_bf16_forward_ws(fake, 16, 64, 0, 64, 3, fake, None, 64, fake, 64, 0, None, 1, 0, 0,
                                           fake, 16, None)
    assert rc == -3 and "workspace" in lib.lgdhip_last_error().decode()


def test_jacobi_zrange_rejects_bad_ranges_without_gpu():
    lib = lgdhip.load()
    g = lgdhip.make_grid(3, 16, 16, 12, halo=1, elem_bytes=8)
    fake = C.c_void_p(0x1000), C.c_void_p(0x2000)
    k = C.c_int(0)
    for z0, z1 in ((0, 0), (5, 3), (-1, 4), (0, 13)):
        rc = lib.lgdhip_jacobi3d_f64_run_zrange(fake[0], fake[1], C.byref(g), z0, z1, None, C.byref(k), None)
        assert rc == -1 and "z0" in lib.lgdhip_last_error().decode()
    t = lgdhip.tile(variant=2)  # a single-sweep variant has no z range
    rc = lib.lgdhip_jacobi3d_f64_run_zrange(fake[0], fake[1], C.byref(g), 0, 4, C.byref(t), C.byref(k), None)
    assert rc == -2 and "temporal-blocking" in lib.lgdhip_last_error().decode()


def test_jacobi_launch_plan_without_gpu():
    """lgdhip_jacobi3d_f64_plan (host-only): the default picks 4 sweeps per
    launch with two rows per thread on a 128^3 cube (13 chunks of 10
    planes: 234 work-groups, one per CU) and on 512^3 (64-plane chunks), one
    row per thread on 64^3 (its 2 x 8 tiles need 4-plane chunks to give the
    CUs 256 work-groups); an explicit variant
    and z chunk are kept; a view falls back to single sweeps and rejects
    temporal blocking, as lgdhip_jacobi3d_f64_run does."""
    lib = lgdhip.load()

    def plan(g, tile=None):
        v, z = C.c_int(-1), C.c_int(-1)
        rc = lib.lgdhip_jacobi3d_f64_plan(C.byref(g), C.byref(tile) if tile else None, C.byref(v), C.byref(z))
        return rc, v.value, z.value

    assert plan(lgdhip.make_grid(3, 128, 128, 128, halo=1, elem_bytes=8)) == (0, 10, 10)
    assert plan(lgdhip.make_grid(3, 64, 64, 64, halo=1, elem_bytes=8)) == (0, 6, 4)
    assert plan(lgdhip.make_grid(3, 512, 512, 512, halo=1, elem_bytes=8)) == (0, 10, 64)
    g = lgdhip.make_grid(3, 96, 70, 40, halo=1, elem_bytes=8)
    assert plan(g, lgdhip.tile(variant=9, bz=7)) == (0, 9, 7)
    v = lgdhip.lgdhip_grid()
    lgdhip.call("lgdhip_grid_view", C.byref(g), 1, 1, 1, 20, 20, 20, C.byref(v))
    assert plan(v)[:2] == (0, 2)
    assert plan(v, lgdhip.tile(variant=10))[0] == -2


def test_poisson_sweep_plan_without_gpu():
    """lgdhip_poisson_sweep_plan (host-only): the default takes three rows per
    thread (variant 6) for launches of at least 2048 of its tiles (the
    512-problem ensemble) and one row (variant 3) for small sets (16 problems
    of 128^3: 384 tiles); requested variants are kept."""
    lib = lgdhip.load()

    def plan(sizes, requested=0):
        probs = (lgdhip.lgdhip_poisson_problem * len(sizes))()
        for k, n in enumerate(sizes):
            probs[k].n = n
        lib.lgdhip_poisson_layout(probs, len(sizes))
        v = C.c_int(-1)
        assert lib.lgdhip_poisson_sweep_plan(probs, len(sizes), requested, C.byref(v)) == 0
        return v.value

    ensemble = [32 + 16 * (k % 11) for k in range(512)]
    assert plan(ensemble) == 6
    assert plan([128] * 16) == 3
    assert plan([128] * 16, 5) == 5 and plan(ensemble, 3) == 3


def test_new_entry_points_reject_bad_arguments_without_gpu():
    """Argument checks that return before any device work: the plan queries
    need their output pointers and valid problems; weighted sweeps need
    omega in (0, 1]."""
    lib = lgdhip.load()
    g = lgdhip.make_grid(3, 32, 32, 32, halo=1, elem_bytes=8)
    z = C.c_int()
    assert lib.lgdhip_jacobi3d_f64_plan(C.byref(g), None, None, C.byref(z)) == -1
    assert "NULL" in lib.lgdhip_last_error().decode()
    probs = (lgdhip.lgdhip_poisson_problem * 2)()
    probs[0].n, probs[1].n = 16, 0
    v = C.c_int()
    assert lib.lgdhip_poisson_sweep_plan(probs, 2, 0, C.byref(v)) == -1      # n = 0
    assert lib.lgdhip_poisson_sweep_plan(probs, 0, 0, C.byref(v)) == -1      # no problems
    probs[1].n = 8
    lib.lgdhip_poisson_layout(probs, 2)
    assert lib.lgdhip_poisson_sweep_plan(probs, 2, 0, None) == -1
    fake = [C.c_void_p(0x1000 * (i + 1)) for i in range(4)]
    for omega in (0.0, -0.5, 1.5):
        rc = lib.lgdhip_poisson_damped_sweeps_f32(fake[0], fake[1], fake[2], probs, 2, 4, C.c_float(omega), None,
                                                  fake[3], C.c_size_t(1 << 20), None, None)
        assert rc == -1 and "omega" in lib.lgdhip_last_error().decode()
