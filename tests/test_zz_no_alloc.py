"""The no-allocation Poisson sweep entry point
(lgdhip_poisson_sweeps_f32_noalloc) against lgdhip_poisson_sweeps_f32 on the
same inputs: bit-identical corrections and the same result buffer, over two
calls in a row on one device (the second with a smaller problem set, so the
static scratch still holds the first call's longer tile lists), and the
size limit rejected on the host before any device work."""
import ctypes as C

import numpy as np
import pytest

from libgeodecomp_amd import lgdhip

SCRATCH_BYTES = 8 << 20


def problems_of(sizes):
    probs = (lgdhip.lgdhip_poisson_problem * len(sizes))()
    for k, n in enumerate(sizes):
        probs[k].n = n
    elems = lgdhip.load().lgdhip_poisson_layout(probs, len(sizes))
    return probs, elems


def test_noalloc_rejects_sets_beyond_the_scratch():
    """Sixteen 512^3 problems need ~9.5 MiB of tile lists: refused with the
    workspace error before the device is touched (host buffers as stand-ins
    for the pools are never read)."""
    probs, elems = problems_of([512] * 16)
    dummy = np.zeros(4, np.float32)
    p = dummy.ctypes.data_as(C.c_void_p)
    in_b = C.c_int(-1)
    lib = lgdhip.load()
    rc = lib.lgdhip_poisson_sweeps_f32_noalloc(p, C.c_void_p(dummy.ctypes.data + 8), p, probs, 16, 4, None,
                                               C.byref(in_b), None)
    assert rc == -3  # LGDHIP_ERR_WORKSPACE
    assert "static scratch holds 8388608" in lib.lgdhip_last_error().decode()
    assert in_b.value == -1


def run_pair(hip, sizes, sweeps, variant, rng):
    """Both entry points on the same e, r pools; returns the plan used."""
    probs, elems = problems_of(sizes)
    count = len(sizes)
    plan = C.c_int(0)
    hip.call("lgdhip_poisson_sweep_plan", probs, count, variant, C.byref(plan))
    ws_bytes = hip.load().lgdhip_poisson_workspace_bytes(probs, count)
    ws = hip.DeviceBuffer(ws_bytes)
    e0 = np.zeros(elems, np.float32)
    r0 = np.zeros(elems, np.float32)
    for k, n in enumerate(sizes):
        o, m = probs[k].offset, (n + 2) ** 3
        e = np.zeros((n + 2,) * 3, np.float32)
        r = (10 * rng.standard_normal((n + 2,) * 3)).astype(np.float32)  # noise in r's halo
        e[1:-1, 1:-1, 1:-1] = rng.standard_normal((n, n, n))
        r[1:-1, 1:-1, 1:-1] = rng.standard_normal((n, n, n))
        e0[o:o + m] = e.ravel()
        r0[o:o + m] = r.ravel()
    r = hip.DeviceBuffer(elems * 4)
    r.upload(r0)
    outs = []
    tile = hip.tile(variant=variant)
    for name in ("lgdhip_poisson_sweeps_f32", "lgdhip_poisson_sweeps_f32_noalloc"):
        ea, eb = hip.DeviceBuffer(elems * 4), hip.DeviceBuffer(elems * 4)
        ea.upload(e0)
        eb.upload(np.zeros(elems, np.float32))
        in_b = C.c_int(-1)
        print(f"[case] {name} sizes {sorted(set(sizes))} x{count} sweeps {sweeps} variant {variant}", flush=True)
        if name.endswith("noalloc"):
            hip.call(name, ea.ptr, eb.ptr, r.ptr, probs, count, sweeps, C.byref(tile), C.byref(in_b), None)
        else:
            hip.call(name, ea.ptr, eb.ptr, r.ptr, probs, count, sweeps, C.byref(tile), ws.ptr, ws_bytes,
                     C.byref(in_b), None)
        hip.call("lgdhip_device_synchronize")
        out = np.empty(elems, np.float32)
        (eb if in_b.value else ea).download(out)
        outs.append((in_b.value, out))
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(outs[0][1], outs[1][1])
    # the sweeps did change the correction
    assert not np.array_equal(outs[0][1], e0)
    return plan.value


@pytest.mark.gpu
def test_noalloc_matches_workspace_entry_point_twice(hip):
    """First call: 2000 problems of 32^3 and eight of 48..144 (above the
    three-row kernel's launch threshold of 2048 of its tiles, so variant 6
    with its tile list); second call: a smaller set of odd sizes through the
    one-row blocked kernel plus single sweeps for the remainder."""
    rng = np.random.default_rng(2026)
    big = [32] * 2000 + [48, 64, 80, 96, 112, 128, 144, 144]
    assert run_pair(hip, big, 12, 0, rng) == 6
    small = [1, 5, 17, 33, 70]
    assert run_pair(hip, small, 10, 0, rng) == 3


@pytest.mark.gpu
@pytest.mark.parametrize("variant,sweeps", [(1, 3), (2, 6), (5, 9)])
def test_noalloc_matches_other_variants(hip, variant, sweeps):
    rng = np.random.default_rng(7 + variant)
    run_pair(hip, [8, 33, 64, 97], sweeps, variant, rng)
