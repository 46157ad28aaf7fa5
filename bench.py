#!/usr/bin/env python3
"""Benchmark of the HipSimulator sweep path on one MI355X.

Default workload (the configuration the headline GLUPS figure is quoted on):
Conway's Game of Life, 16384 x 16384 uint8 cells, Bernoulli(0.3) from seed,
constant dead edge (Cube topology), 1000 generations.  One benchmark step is
one whole run: restore the pristine uint8 input (device copy), pack, 1000
generations, unpack -- uint8 grid in HBM in, uint8 grid in HBM out.

Other workloads (`--workload jacobi|lbm|poisson|learned`) measure the
Jacobi3D FP64 and D3Q19 FP32 families on their configured shapes, the
batched Poisson ensemble (one step = the whole ensemble solved to 1e-6) and
the learned initial guess (set-up = data generation and training; one step =
held-out problems solved from the zero and from the learned guess; value =
the wall-time ratio).  Run those two with `--steps 1 --warmup 0`.

Prints ONE JSON line (see the keys below).  GPU time comes from HIP events
recorded on the stream the kernels run on; the dominant kernel's average
launch duration is measured the same way over a block of launches.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

from libgeodecomp_amd import lgdhip, synthetic  # noqa: E402

METRIC = "GLUPS + % HBM roofline; learned vs plain Jacobi time to 1e-6 residual"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def timed(stream, fn, reps=1):
    """Run fn() `reps` times between two events on `stream`; return ms."""
    e0, e1 = lgdhip.Event(), lgdhip.Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.sync()
    return e0.elapsed_ms(e1)


def load_traffic(key: str):
    """HBM bytes per launch of a dominant kernel on one workload shape
    ("kernel@shape") from the committed rocprof PMC summaries
    (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE scaled by the copy
    calibration), or None when that shape was not profiled."""
    p = REPO / "profiles" / "traffic.json"
    if not p.exists():
        return None
    try:
        entry = json.loads(p.read_text()).get(key)
    except ValueError:
        return None
    return entry.get("hbm_bytes_per_launch") if entry else None


def cpu_threads() -> int:
    n = os.cpu_count() or 1
    return max(1, min(16, n))


def load_oracle():
    """The CPU oracle, timed as the CPU baseline (never as the product)."""
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_lib  # noqa: E402

    oracle_lib.load()
    return oracle_lib


# ----------------------------------------------------------------------
# Game of Life 16384^2 x 1000
# ----------------------------------------------------------------------
class Life:
    name = "life"
    dtype = "u8"
    bytes_per_lup = 2.0  # 1 B read + 1 B write of the uint8 cell format

    def __init__(self, n=16384, generations=1000, seed=synthetic.DEFAULT_SEED, per_launch=31, variant=7):
        self.n, self.gens, self.seed = n, generations, seed
        self.tile = lgdhip.tile(steps_per_launch=per_launch, variant=variant)
        self.per_launch = per_launch
        self.stream = lgdhip.Stream()
        self.g = lgdhip.make_grid(2, n, n, 1, halo=1, elem_bytes=1)
        t0 = time.time()
        host = synthetic.life_initial(n, n, 0.3, seed)
        log(f"life input {n}x{n} generated in {time.time() - t0:.1f}s, density {host.mean():.4f}")
        self.pristine = lgdhip.DeviceGrid(self.g, np.uint8)
        self.pristine.fill_edge(0, self.stream.h)
        self.pristine.load(host, self.stream.h)
        self.work = lgdhip.DeviceGrid(self.g, np.uint8)
        self.nbytes = lgdhip.grid_bytes(self.g)
        self.ws_bytes = lgdhip.load().lgdhip_life2d_workspace_bytes(C.byref(self.g))
        self.ws = lgdhip.DeviceBuffer(self.ws_bytes)
        self.host = host

    def config(self):
        return {"workload": f"Conway Game of Life 2D uint8 {self.n}x{self.n}, Bernoulli(0.3), dead Cube edge, "
                            f"{self.gens} generations", "grid": [self.n, self.n], "generations": self.gens,
                "seed": self.seed, "generations_per_launch": self.per_launch, "parallelism": "single GPU"}

    def lups_per_step(self):
        return float(self.n) * self.n * self.gens

    def step(self):
        s = self.stream.h
        lgdhip.call("lgdhip_memcpy_d2d", self.work.ptr, self.pristine.ptr, self.nbytes, s)
        lgdhip.call("lgdhip_life2d_run", self.work.ptr, C.byref(self.g), 0, self.gens, self.ws.ptr,
                    self.ws_bytes, C.byref(self.tile), s)

    def dominant_kernel(self):
        """(kernel name, avg ms per launch, algorithmic bytes per launch)."""
        s = self.stream.h
        half = self.ws_bytes // 2
        pa, pb = self.ws.value, self.ws.value + half
        lgdhip.call("lgdhip_life2d_pack", self.pristine.ptr, C.byref(self.g), C.c_void_p(pa), s)
        launches = 64
        gens = launches * self.per_launch
        in_b = C.c_int()
        ms = timed(self.stream, lambda: lgdhip.call("lgdhip_life2d_packed_run", C.c_void_p(pa), C.c_void_p(pb),
                                                    C.byref(self.g), 0, gens, C.byref(self.tile),
                                                    C.byref(in_b), s))
        per_launch_bytes = self.bytes_per_lup * self.n * self.n * self.per_launch
        # variant 0 is the library default (a stacked-wave kernel)
        name = "life_packed_tb" if 1 <= self.tile.variant <= 4 else "life_packed_stacked"
        self.shape = f"{self.n}x{self.n},{self.per_launch}gen"
        return name, ms / launches, per_launch_bytes

    def cpu_baseline(self, budget_s=15.0):
        oracle = load_oracle()
        th = cpu_threads()
        sample = self.host[:4096, :4096].copy()
        gens = 8
        t0 = time.perf_counter()
        oracle.life2d(sample, gens, 0, th)
        dt = time.perf_counter() - t0
        # scale the sample to roughly the time budget
        gens = max(8, int(gens * budget_s / max(dt, 1e-3)))
        gens = min(gens, 2000)
        t0 = time.perf_counter()
        oracle.life2d(sample, gens, 0, th)
        dt = time.perf_counter() - t0
        return {"value": 1e-9 * sample.size * gens / dt, "unit": "GLUPS", "cores": th, "kind": "port",
                "sample": f"C oracle (OpenMP, {th} threads) on the top-left 4096x4096 of the same input, "
                          f"{gens} generations, {dt:.1f}s"}


# ----------------------------------------------------------------------
# Jacobi3D FP64
# ----------------------------------------------------------------------
class Jacobi:
    name = "jacobi"
    dtype = "f64"
    bytes_per_lup = 16.0

    def __init__(self, n=128, steps=100, seed=synthetic.DEFAULT_SEED, variant=0, by=0, bz=0, flags=0, **_):
        self.n, self.sweeps, self.seed = n, steps, seed
        self.tile = lgdhip.tile(by=by, bz=bz, variant=variant, flags=flags)
        self.stream = lgdhip.Stream()
        self.g = lgdhip.make_grid(3, n, n, n, halo=1, elem_bytes=8)
        host = synthetic.jacobi_initial(n, n, n, seed)
        self.pristine = lgdhip.DeviceGrid(self.g, np.float64)
        self.a = lgdhip.DeviceGrid(self.g, np.float64)
        self.b = lgdhip.DeviceGrid(self.g, np.float64)
        for d in (self.pristine, self.a, self.b):
            d.fill_edge(0.0, self.stream.h)
        self.pristine.load(host, self.stream.h)
        self.nbytes = lgdhip.grid_bytes(self.g)
        self.host = host

    def config(self):
        c = {"workload": f"Jacobi3D 7-point FP64 {self.n}^3, zero Cube edge, U[0,1) interior, {self.sweeps} steps",
             "grid": [self.n] * 3, "steps": self.sweeps, "seed": self.seed, "parallelism": "single GPU"}
        if getattr(self, "plan", None):
            c["launch"] = self.plan  # the library's plan: kernel variant and z planes per work-group
        return c

    def lups_per_step(self):
        return float(self.n) ** 3 * self.sweeps

    def step(self):
        s = self.stream.h
        lgdhip.call("lgdhip_memcpy_d2d", self.a.ptr, self.pristine.ptr, self.nbytes, s)
        lgdhip.call("lgdhip_jacobi3d_f64_run", self.a.ptr, self.b.ptr, C.byref(self.g), self.sweeps,
                    C.byref(self.tile), None, s)

    # sweeps per launch of the temporal-blocking variants (9 and 10 hold two
    # rows per thread); the library resolves variant 0 by its launch plan
    TB_DEPTH = {4: 2, 5: 3, 6: 4, 7: 2, 8: 3, 9: 4, 10: 4}

    def dominant_kernel(self):
        v, zc = C.c_int(0), C.c_int(0)
        lgdhip.call("lgdhip_jacobi3d_f64_plan", C.byref(self.g), C.byref(self.tile), C.byref(v), C.byref(zc))
        variant = v.value
        k = self.TB_DEPTH.get(variant, 1)
        sweeps = 240   # a multiple of every blocking depth: launches of one kernel only
        ms = timed(self.stream, lambda: lgdhip.call("lgdhip_jacobi3d_f64_run", self.a.ptr, self.b.ptr,
                                                    C.byref(self.g), sweeps, C.byref(self.tile), None,
                                                    self.stream.h))
        if k > 1:
            name = "jacobi3d_f64_tb2" if variant in (9, 10) else "jacobi3d_f64_tb"
        else:
            name = "jacobi3d_f64_zmarch" if self.tile.variant == 1 else "jacobi3d_f64_zbatch"
        # one launch advances k sweeps: algorithmic bytes are k sweeps' worth
        self.shape = f"{self.n}^3,K{k}" if k > 1 else f"{self.n}^3"
        self.plan = {"variant": variant, "zchunk": zc.value}
        return name, ms / (sweeps // k), k * self.bytes_per_lup * float(self.n) ** 3

    def cpu_baseline(self, budget_s=15.0):
        oracle = load_oracle()
        th = cpu_threads()
        steps = 64  # a probe long enough that thread start-up does not dominate the rate
        t0 = time.perf_counter()
        oracle.jacobi3d(self.host, steps, 0.0, th)
        dt = time.perf_counter() - t0
        steps = max(4, min(200000, int(steps * budget_s / max(dt, 1e-3))))
        t0 = time.perf_counter()
        oracle.jacobi3d(self.host, steps, 0.0, th)
        dt = time.perf_counter() - t0
        return {"value": 1e-9 * self.host.size * steps / dt, "unit": "GLUPS", "cores": th, "kind": "port",
                "sample": f"C oracle (OpenMP, {th} threads), same {self.n}^3 input, {steps} steps, {dt:.1f}s"}


# ----------------------------------------------------------------------
# D3Q19 LBM FP32
# ----------------------------------------------------------------------
class LBM:
    name = "lbm"
    dtype = "f32"
    bytes_per_lup = 153.0  # 19 x 4 B read + 19 x 4 B write + 1 B wall flags

    def __init__(self, n=256, steps=2000, seed=synthetic.DEFAULT_SEED, by=0, flags=0, variant=0, bx=0, bz=0,
                 layout=lgdhip.LAYOUT_Z_SLABS, **_):
        self.n, self.sweeps, self.seed = n, steps, seed
        self.tile = lgdhip.tile(bx=bx, by=by, bz=bz, flags=flags, variant=variant)
        self.stream = lgdhip.Stream()
        s = self.stream.h
        self.g = lgdhip.make_grid(3, n, n, n, halo=1, elem_bytes=4, members=19, layout=layout)
        self.fg = lgdhip.make_grid(3, n, n, n, halo=1, elem_bytes=1, members=1)
        sys.path.insert(0, str(REPO / "tests"))
        ux, uy, uz = synthetic.lbm_initial_velocity(n, n, n, 1e-3, seed)
        f0 = equilibrium_f32(np.ones_like(ux), ux, uy, uz)  # shifted populations
        self.pristine = lgdhip.DeviceGrid(self.g, np.float32)
        self.a = lgdhip.DeviceGrid(self.g, np.float32)
        self.b = lgdhip.DeviceGrid(self.g, np.float32)
        for d in (self.pristine, self.a, self.b):
            d.fill_edge(np.zeros(19, np.float32), s)
        self.pristine.load(f0, s)
        del f0
        self.flags = lgdhip.DeviceGrid(self.fg, np.uint8)
        self.flags.fill_edge(0, s)
        lgdhip.call("lgdhip_lbm_cavity_flags", self.flags.ptr, C.byref(self.fg), s)
        self.p = lgdhip.lgdhip_lbm_params(1.7, 0.05, 0.0, 0.0)
        self.nbytes = lgdhip.grid_bytes(self.g)
        self.ux = ux

    def config(self):
        return {"workload": f"LBM D3Q19 BGK lid-driven cavity FP32 {self.n}^3, omega 1.7, lid 0.05, "
                            f"{self.sweeps} steps", "grid": [self.n] * 3, "steps": self.sweeps, "seed": self.seed,
                "layout": "z slabs" if self.g.layout == lgdhip.LAYOUT_Z_SLABS else "member planes",
                "parallelism": "single GPU"}

    def lups_per_step(self):
        return float(self.n) ** 3 * self.sweeps

    def step(self):
        s = self.stream.h
        lgdhip.call("lgdhip_memcpy_d2d", self.a.ptr, self.pristine.ptr, self.nbytes, s)
        lgdhip.call("lgdhip_lbm_d3q19_f32_run", self.a.ptr, self.b.ptr, self.flags.ptr, C.byref(self.g),
                    C.byref(self.p), self.sweeps, C.byref(self.tile), None, s)

    def dominant_kernel(self):
        # variant 3 advances two steps per launch (credited two steps' bytes)
        per = 2 if self.tile.variant == 3 else 1
        launches = 20
        ms = timed(self.stream, lambda: lgdhip.call("lgdhip_lbm_d3q19_f32_run", self.a.ptr, self.b.ptr,
                                                    self.flags.ptr, C.byref(self.g), C.byref(self.p),
                                                    launches * per, C.byref(self.tile), None, self.stream.h))
        self.shape = f"{self.n}^3" + (",zslabs" if self.g.layout == lgdhip.LAYOUT_Z_SLABS else "") + \
            (",2steps" if per == 2 else "")
        name = "lbm_d3q19_pull2" if per == 2 else "lbm_d3q19_pull"
        return name, ms / launches, per * self.bytes_per_lup * float(self.n) ** 3

    def cpu_baseline(self, budget_s=15.0):
        oracle = load_oracle()
        th = cpu_threads()
        m = 64
        ux, uy, uz = synthetic.lbm_initial_velocity(m, m, m, 1e-3, self.seed)
        f0 = oracle.lbm_equilibrium(np.ones_like(ux), ux, uy, uz)
        fl = oracle.lbm_cavity_flags(m, m, m)
        steps = 2
        t0 = time.perf_counter()
        oracle.lbm_run(f0, fl, steps, 1.7, (0.05, 0.0, 0.0), th)
        dt = time.perf_counter() - t0
        steps = max(2, min(2000, int(steps * budget_s / max(dt, 1e-3))))
        t0 = time.perf_counter()
        oracle.lbm_run(f0, fl, steps, 1.7, (0.05, 0.0, 0.0), th)
        dt = time.perf_counter() - t0
        return {"value": 1e-9 * m ** 3 * steps / dt, "unit": "GLUPS", "cores": th, "kind": "port",
                "sample": f"C oracle FP64 (OpenMP, {th} threads), 64^3 cavity, {steps} steps, {dt:.1f}s"}


def equilibrium_f32(rho, jx, jy, jz):
    """Incompressible D3Q19 equilibrium (same formula as the cell) in numpy,
    returned in the device's shifted storage g = f_eq - w (FP64 math, FP32
    result)."""
    ex = [0, 0, 1, -1, 0, 0, 0, -1, -1, 1, 1, -1, -1, 1, 1, 0, 0, 0, 0]
    ey = [0, 1, 0, 0, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1]
    ez = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1]
    w = [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12
    d = rho / 3.0 - 0.5 * (jx * jx + jy * jy + jz * jz)
    f = np.empty((19,) + rho.shape, dtype=np.float32)
    for i in range(19):
        eu = ex[i] * jx + ey[i] * jy + ez[i] * jz
        f[i] = 3.0 * w[i] * (d + eu + 1.5 * eu * eu) - w[i]
    return f


def poisson_sweep_kernel(stream, e0, e1, f, probs, count, ws, ws_bytes, cells, wl, shape, variant=0):
    """(name, ms per launch, algorithmic bytes per launch) of the Poisson
    solver's sweep kernel: temporal blocking, 4 sweeps per launch (12 B per
    sweep-LUP, so 4 x 12 B x cells per launch); variant 0 is the library
    default (6, three rows per thread, on large launches, else 3, one row)."""
    k, sweeps = 4, 48
    in_b = C.c_int(0)
    tile = lgdhip.tile(variant=variant)
    ms = timed(stream, lambda: lgdhip.call("lgdhip_poisson_sweeps_f32", e0.ptr, e1.ptr, f.ptr, probs, count, sweeps,
                                           C.byref(tile), ws.ptr, ws_bytes, C.byref(in_b), stream.h))
    wl.shape = shape
    planned = C.c_int(0)  # the library resolves the default by launch size
    lgdhip.call("lgdhip_poisson_sweep_plan", probs, count, variant, C.byref(planned))
    wl.sweep_plan = planned.value
    return ("poisson_sweep_tb2" if planned.value in (4, 5, 6) else "poisson_sweep_tb"), ms / (sweeps // k), \
        k * 12.0 * cells


# ----------------------------------------------------------------------
# Batched Poisson ensemble (FP32 sweeps, FP64 residuals)
# ----------------------------------------------------------------------
class Poisson:
    """512 independent n^3 problems, n uniform in {32, 48, ..., 192}, 1-8
    Gaussian sources, Dirichlet faces U[-1,1]; one benchmark step solves the
    whole ensemble from the zero guess to a 1e-6 relative FP64 residual
    (one batched solve, hipGraph chunks).  GLUPS counts sweep-cell updates
    (sum over problems of n^3 x sweeps)."""
    name = "poisson"
    dtype = "f32"   # the sweeps; residuals and updates are FP64
    bytes_per_lup = 12.0  # read e + read r + write e (FP32)

    def __init__(self, n=512, steps=0, seed=synthetic.DEFAULT_SEED, use_graph=1, tol=1e-6, variant=0, **_):
        self.count, self.seed = n, seed   # --size overrides the problem count
        self.use_graph = use_graph
        self.tol = tol
        self.variant = variant            # sweep variant (0: the library default)
        self.stream = lgdhip.Stream()
        s = self.stream.h
        t0 = time.perf_counter()
        problems = synthetic.poisson_ensemble(self.count, seed=seed)
        log(f"generated {self.count} problems on the host in {time.perf_counter() - t0:.1f} s")
        self.sizes = np.array([p[0] for p in problems], dtype=np.int64)
        self.probs = (lgdhip.lgdhip_poisson_problem * self.count)()
        for k, (m, faces, _) in enumerate(problems):
            self.probs[k].n = m
            for j in range(6):
                self.probs[k].face[j] = faces[j]
        lib = lgdhip.load()
        self.elems = lib.lgdhip_poisson_layout(self.probs, self.count)
        self.ws_bytes = lib.lgdhip_poisson_workspace_bytes(self.probs, self.count)
        self.ws = lgdhip.DeviceBuffer(self.ws_bytes)
        self.u = lgdhip.DeviceBuffer(self.elems * 8)
        self.f = lgdhip.DeviceBuffer(self.elems * 4)
        for k, (m, _, f) in enumerate(problems):
            host = np.ascontiguousarray(f, dtype=np.float32).ravel()
            dst = C.c_void_p(self.f.value + self.probs[k].offset * 4)
            lgdhip.call("lgdhip_memcpy_h2d", dst, host.ctypes.data_as(C.c_void_p), host.nbytes, s)
        self.sample = problems[int(np.argmin(np.abs(self.sizes - 64)))]
        del problems
        self.cells = float(np.sum(self.sizes.astype(np.float64) ** 3))
        self.last_lups = 0.0
        self.last = None

    def config(self):
        sw = getattr(self, "total_sweeps", None)
        return {"workload": f"batched Poisson ensemble, {self.count} problems n in {{32..192 step 16}}, 1-8 Gaussian "
                            f"sources, Dirichlet faces U[-1,1], zero guess to {self.tol:g} relative FP64 residual",
                "problems": self.count, "cells": int(self.cells), "seed": self.seed,
                "sweeps_max": int(sw.max()) if sw is not None else None,
                "sweep_lups": self.last_lups,
                "converged": int(sum(r.converged for r in self.last)) if self.last is not None else None,
                "hip_graph": bool(self.use_graph), "sweep_variant": self.variant, "parallelism": "single GPU"}

    def lups_per_step(self):
        return self.last_lups

    SWEEPS_PER_CALL = 10000   # bounded solver calls, so progress is reported between them

    def step(self):
        """Zero guess -> 1e-6 for the whole ensemble.  The solve is issued as
        calls of at most SWEEPS_PER_CALL sweeps per problem, each continuing
        from the FP64 state the previous one left, all against the zero
        guess's residual norms (the same iteration as one unbounded call,
        with one extra FP64 residual per call)."""
        s = self.stream.h
        lgdhip.call("lgdhip_fill_u8", self.u.ptr, 0, self.elems * 8, s)
        lgdhip.call("lgdhip_poisson_set_faces", self.u.ptr, self.probs, self.count, self.ws.ptr, self.ws_bytes, s)
        ref = (C.c_double * self.count)()
        lgdhip.call("lgdhip_poisson_residual_norms", self.u.ptr, self.f.ptr, self.probs, self.count, self.ws.ptr,
                    self.ws_bytes, ref, s)
        opt = lgdhip.lgdhip_poisson_options()
        opt.rel_tol, opt.chunk, opt.max_sweeps, opt.use_graph = self.tol, 64, self.SWEEPS_PER_CALL, self.use_graph
        opt.sweep_variant = self.variant
        opt.reference_norms = ref
        total = np.zeros(self.count, dtype=np.int64)
        res = (lgdhip.lgdhip_poisson_result * self.count)()
        for call in range(1000):
            lgdhip.call("lgdhip_poisson_batch_solve", self.u.ptr, self.f.ptr, self.probs, self.count, C.byref(opt),
                        self.ws.ptr, self.ws_bytes, res, s)
            total += np.array([r.sweeps for r in res], dtype=np.int64)
            done = sum(r.converged for r in res)
            log(f"solver call {call}: {done}/{self.count} converged, max sweeps so far {total.max()}")
            if done == self.count:
                break
        self.last = res
        self.total_sweeps = total
        self.last_lups = float(np.sum(self.sizes.astype(np.float64) ** 3 * total))

    def dominant_kernel(self):
        """The solver's sweep kernel over every problem (the library default,
        4 sweeps per launch)."""
        e0 = lgdhip.DeviceBuffer(self.elems * 4)
        e1 = lgdhip.DeviceBuffer(self.elems * 4)
        for b in (e0, e1):
            lgdhip.call("lgdhip_fill_u8", b.ptr, 0, self.elems * 4, self.stream.h)
        return poisson_sweep_kernel(self.stream, e0, e1, self.f, self.probs, self.count, self.ws, self.ws_bytes,
                                    self.cells, self, f"{self.count}problems", self.variant)

    def cpu_baseline(self, budget_s=15.0):
        oracle = load_oracle()
        m, faces, f = self.sample
        u = np.zeros((m + 2,) * 3)
        oracle.poisson_faces(u, faces)
        sweeps = 20
        t0 = time.perf_counter()
        oracle.poisson_jacobi(u.copy(), f, 1e-300, max_sweeps=sweeps, check=sweeps)
        dt = time.perf_counter() - t0
        sweeps = max(20, min(200000, int(sweeps * budget_s / max(dt, 1e-3))))
        t0 = time.perf_counter()
        oracle.poisson_jacobi(u.copy(), f, 1e-300, max_sweeps=sweeps, check=sweeps)
        dt = time.perf_counter() - t0
        return {"value": 1e-9 * m ** 3 * sweeps / dt, "unit": "GLUPS", "cores": 1, "kind": "port",
                "sample": f"C oracle plain FP64 Jacobi (1 thread), one {m}^3 problem of the ensemble, {sweeps} "
                          f"sweeps, {dt:.1f}s"}


# ----------------------------------------------------------------------
# Learned initial guess (U-Net on a coarse-grid base) vs the zero guess
# ----------------------------------------------------------------------
class Learned:
    """Set-up (untimed, reported in config): solve `train` 128^3 problems with
    the batched Jacobi solver, solve their 31^3 coarse problems, train the
    U-Net for `steps` Adam steps.  One benchmark step: the held-out problems
    solved to 1e-6 from the zero guess and from the learned guess (coarse
    solve + prolongation + network + polishing), same solver and reference
    norms.  value = zero-guess wall time / learned wall time."""
    name = "learned"
    dtype = "bf16"   # the network (BF16 operands, FP32 accumulation); polishing is FP32/FP64
    bytes_per_lup = 12.0

    def __init__(self, n=128, steps=20000, seed=synthetic.DEFAULT_SEED, train=512, eval=16, lr=5e-4, **_):
        from libgeodecomp_amd import learned_guess as lgm
        self.lgm = lgm
        self.n, self.train_steps, self.seed, self.ntrain, self.neval = n, steps, seed, train, eval
        self.lr = lr
        self.m = lgm.default_coarse_size(n)
        self.stream = lgdhip.Stream()
        t0 = time.perf_counter()
        data = lgm.Ensemble(synthetic.poisson_ensemble(train, sizes=(n,), seed=seed), stream=self.stream.h)
        res = data.solve(1e-6)
        base = lgm.CoarseBase(data, self.m)
        base.compute()
        lgdhip.call("lgdhip_device_synchronize")
        self.generate_s = time.perf_counter() - t0
        self.gen_sweeps = float(np.mean([r.sweeps for r in res]))
        log(f"generated {train} solutions of {n}^3 in {self.generate_s:.1f} s")
        self.model = lgm.LearnedGuess(n, seed=seed % 1000, stream=self.stream.h)
        t0 = time.perf_counter()
        self.log = self.model.train(data, base, steps, lr=lr, log_every=1000, verbose=True)
        lgdhip.call("lgdhip_device_synchronize")
        self.train_s = time.perf_counter() - t0
        del base, data
        self.eval = lgm.Ensemble(synthetic.poisson_ensemble(eval, sizes=(n,), seed=seed + 7919), stream=self.stream.h)
        self.result = None

    def config(self):
        r = self.result or {}
        return {"workload": f"learned initial guess vs zero guess, {self.neval} held-out Poisson problems {self.n}^3 "
                            f"(Gaussian sources, Dirichlet faces), Jacobi polishing to 1e-6 relative FP64 residual",
                "network": "3D U-Net 16-32-64-128, 3x3x3 BF16 convs on MFMA, FP32 accumulation, bias-free",
                "coarse_grid": self.m, "train_problems": self.ntrain, "train_steps": self.train_steps,
                "peak_lr": self.lr,
                "generate_s": round(self.generate_s, 1), "train_s": round(self.train_s, 1),
                "final_loss": self.log[-1][1] if self.log else None, "arms": r.get("arms"),
                "sweep_ratio": r.get("sweep_ratio"), "seed": self.seed, "parallelism": "single GPU"}

    def step(self):
        self.result = self.lgm.evaluate(self.model, self.eval, self.m)
        for a in self.result["arms"]:
            log(f"{a['arm']:8s} sweeps {a['sweeps_mean']:.0f} wall {a['wall_s']:.3f} s")

    def value(self):
        """(value, unit, ms per step): wall-time speedup of the learned guess."""
        arms = {a["arm"]: a for a in self.result["arms"]}
        return self.result["wall_speedup"], "x (zero-guess wall time / learned-guess wall time)", \
            1e3 * arms["learned"]["wall_s"]

    def lups_per_step(self):
        return 0.0

    def dominant_kernel(self):
        """The polishing sweep over the held-out set (poisson_sweep_f32)."""
        e0 = lgdhip.DeviceBuffer(self.eval.elems * 4)
        e1 = lgdhip.DeviceBuffer(self.eval.elems * 4)
        for b in (e0, e1):
            lgdhip.call("lgdhip_fill_u8", b.ptr, 0, self.eval.elems * 4, self.stream.h)
        return poisson_sweep_kernel(self.stream, e0, e1, self.eval.f, self.eval.probs, self.eval.count, self.eval.ws,
                                    self.eval.ws_bytes, self.neval * float(self.n) ** 3, self,
                                    f"{self.neval}x{self.n}^3")

    def cpu_baseline(self, budget_s=15.0):
        """The oracle's plain FP64 Jacobi on one held-out problem (sweep rate)."""
        oracle = load_oracle()
        u = np.zeros((self.n + 2,) * 3)
        (_, faces, f), = synthetic.poisson_ensemble(1, sizes=(self.n,), seed=self.seed + 7919)
        oracle.poisson_faces(u, faces)
        sweeps = 4
        t0 = time.perf_counter()
        oracle.poisson_jacobi(u.copy(), f, 1e-300, max_sweeps=sweeps, check=sweeps)
        dt = time.perf_counter() - t0
        sweeps = max(4, min(100000, int(sweeps * budget_s / max(dt, 1e-3))))
        t0 = time.perf_counter()
        oracle.poisson_jacobi(u.copy(), f, 1e-300, max_sweeps=sweeps, check=sweeps)
        dt = time.perf_counter() - t0
        return {"value": 1e-9 * self.n ** 3 * sweeps / dt, "unit": "GLUPS", "cores": 1, "kind": "port",
                "sample": f"C oracle plain FP64 Jacobi (1 thread), one held-out {self.n}^3 problem, {sweeps} sweeps, "
                          f"{dt:.1f}s (sweep rate; the zero-guess solve needs ~3e4 sweeps per problem)"}


class Copy:
    """Calibration: the plain 16-byte streaming copy kernel over 2 GiB, the
    measured HBM ceiling the stencil kernels are compared with (not a
    workload of the simulator)."""
    name = "copy"
    dtype = "u8"
    bytes_per_lup = 2.0  # one byte read + one written per copied byte

    def __init__(self, n=1 << 31, **_):
        self.n = n
        self.stream = lgdhip.Stream()
        self.a = lgdhip.DeviceBuffer(n)
        self.b = lgdhip.DeviceBuffer(n)
        lgdhip.call("lgdhip_fill_u8", self.a.ptr, 1, n, self.stream.h)

    def config(self):
        return {"workload": f"device copy calibration, {self.n} bytes", "bytes": self.n, "parallelism": "single GPU"}

    def lups_per_step(self):
        return float(self.n)

    def step(self):
        lgdhip.call("lgdhip_copy_kernel", self.b.ptr, self.a.ptr, self.n, self.stream.h)

    def dominant_kernel(self):
        launches = 10
        ms = timed(self.stream, self.step, reps=launches)
        self.shape = f"{self.n}B"
        return "lgd_copy16_kernel", ms / launches, 2.0 * self.n

    def cpu_baseline(self, budget_s=0.0):
        return None


WORKLOADS = {"life": Life, "jacobi": Jacobi, "lbm": LBM, "poisson": Poisson, "learned": Learned, "copy": Copy}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="life")
    ap.add_argument("--size", type=int, default=0, help="override the grid edge length")
    ap.add_argument("--iters", type=int, default=0, help="override generations / steps per run")
    ap.add_argument("--per-launch", type=int, default=31, help="life: generations per launch")
    ap.add_argument("--variant", type=int, default=-1, help="kernel variant id (-1: tuned default, 0: library default)")
    ap.add_argument("--by", type=int, default=0, help="jacobi/lbm: rows per work-group (0: default)")
    ap.add_argument("--bx", type=int, default=0, help="lbm: work-group width 64 or 256 (0: default)")
    ap.add_argument("--layout", choices=("planes", "zslabs"), default=None,
                    help="lbm: member layout of the populations (default: the tuned one)")
    ap.add_argument("--bz", type=int, default=0, help="jacobi, lbm variant 3: z planes per work-group (0: default)")
    ap.add_argument("--no-xcd-swizzle", action="store_true", help="jacobi/lbm: map work-groups to tiles in launch order")
    ap.add_argument("--xcd-swizzle", action="store_true", help="jacobi/lbm: one contiguous tile range per XCD")
    ap.add_argument("--xcd-rows", action="store_true", help="lbm: x rows of tiles round-robin over the XCDs")
    ap.add_argument("--no-graph", action="store_true", help="poisson: plain launches instead of hipGraph chunks")
    ap.add_argument("--tol", type=float, default=1e-6, help="poisson: relative residual to solve to")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path; the kernels shard as "
              "z slabs with K-plane halos exchanged every K steps (Poisson problems: no exchange at all)",
              file=sys.stderr)
        sys.exit(2)

    lgdhip.load()
    if lgdhip.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    lgdhip.call("lgdhip_set_device", 0)
    log("device:", lgdhip.device_name(0))

    kw = {}
    if args.size:
        kw["n"] = args.size
    if args.iters:
        kw["generations" if args.workload == "life" else "steps"] = args.iters
    if args.workload == "life":
        kw.update(per_launch=args.per_launch, variant=7 if args.variant < 0 else args.variant)
    flags = (lgdhip.TILE_NO_XCD_SWIZZLE if args.no_xcd_swizzle else 0) | \
        (lgdhip.TILE_XCD_SWIZZLE if args.xcd_swizzle else 0) | (lgdhip.TILE_XCD_ROWS if args.xcd_rows else 0)
    if args.workload == "jacobi":
        kw.update(variant=max(0, args.variant), by=args.by, bz=args.bz, flags=flags)
    if args.workload == "lbm":
        kw.update(by=args.by, flags=flags, variant=max(0, args.variant), bx=args.bx, bz=args.bz)
        if args.layout:
            kw.update(layout=lgdhip.LAYOUT_Z_SLABS if args.layout == "zslabs" else lgdhip.LAYOUT_MEMBER_PLANES)
    if args.workload == "poisson":
        kw.update(use_graph=0 if args.no_graph else 1, tol=args.tol, variant=max(0, args.variant))
    wl = WORKLOADS[args.workload](**kw)

    for i in range(args.warmup):
        wl.step()
    wl.stream.sync()
    log(f"warmup {args.warmup} done")

    lgdhip.call("lgdhip_device_synchronize")
    t0 = time.perf_counter()
    ms = timed(wl.stream, wl.step, reps=args.steps)
    lgdhip.call("lgdhip_device_synchronize")
    wall = time.perf_counter() - t0
    ms_per_step = ms / args.steps
    unit = "GLUPS"
    if hasattr(wl, "value"):
        value, unit, ms_per_step = wl.value()
    else:
        value = 1e-9 * wl.lups_per_step() / (ms_per_step * 1e-3)
    log(f"timed {args.steps} steps: {ms:.3f} ms (wall {wall * 1e3:.1f} ms), {value:.3f} {unit}")

    kname, k_ms, k_bytes = wl.dominant_kernel()
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    log(f"dominant kernel {kname}: {k_ms * 1e3:.2f} us/launch, {achieved:.1f} GB/s algorithmic")
    roofline = {"kernel": kname, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": load_traffic(f"{kname}@{getattr(wl, 'shape', '')}"),
                "avg_launch_us": round(k_ms * 1e3, 3), "algorithmic_bytes_per_launch": k_bytes}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = wl.cpu_baseline()
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": wl.dtype,
        "data": "synthetic (SplitMix64 seeded initializer)",
        "config": wl.config(),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
