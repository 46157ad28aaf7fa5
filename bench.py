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


# MI355X published peaks: HBM3E 8 TB/s; VALU issue = 256 CUs x 4 SIMDs, each
# retiring one wave64 instruction per 4 cycles (16 lanes), at the 2.4 GHz
# peak engine clock
VALU_PEAK_GINST = 256 * 4 / 4 * 2.4   # 614.4 G wave64 instructions per second
MFMA_BF16_PEAK_TFLOPS = 2500.0         # dense BF16 matrix cores


def load_counters(key: str):
    """Counter figures per launch of a dominant kernel on one workload shape
    ("kernel@shape") from the committed rocprofv3 PMC summaries
    (profiles/kernel_counters.json): HBM bytes (request-size-weighted
    read/write requests at the memory controllers) and wave64 VALU
    instructions (SQ_INSTS_VALU), or {} when that shape was not profiled."""
    p = REPO / "profiles" / "kernel_counters.json"
    if not p.exists():
        return {}
    try:
        return json.loads(p.read_text()).get(key) or {}
    except ValueError:
        return {}


def roofline(name, shape, ms, alg_bytes, sweeps=1, min_bytes=None):
    """The dominant kernel's roofline from its live average launch time `ms`.

    `alg_bytes` are the algorithmic bytes of the `sweeps` sweeps (or
    generations, or time steps) one launch advances.  A single-sweep kernel
    is held to HBM with achieved = alg_bytes / time.  A temporally blocked
    launch must move the state through HBM only once, in the kernel's own
    storage format (`min_bytes`, default alg_bytes / sweeps), and its sweeps
    are VALU work on chip: both fractions are reported, and the roofline is
    the one the kernel is closer to (the binding one).  The per-sweep rate
    alg_bytes / time goes into `effective_bw` only."""
    t = ms * 1e-3
    c = load_counters(f"{name}@{shape}")
    if min_bytes is None:
        min_bytes = alg_bytes / sweeps
    hbm_alg = min_bytes / t / 1e9
    hbm = {"achieved": round(hbm_alg, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(hbm_alg / HBM_PEAK_GBS, 4), "algorithmic_bytes_per_launch": min_bytes}
    traffic = c.get("hbm_bytes_per_launch")
    if traffic:
        hbm["traffic_GBs"] = round(traffic / t / 1e9, 2)
    out = {"kernel": name, "bound": "hbm", **hbm, "traffic": traffic, "avg_launch_us": round(ms * 1e3, 3),
           "sweeps_per_launch": sweeps, "hbm": dict(hbm)}
    valu = c.get("valu_insts_per_launch")
    if valu:
        rate = valu / t / 1e9
        out["valu"] = {"achieved": round(rate, 2), "peak": VALU_PEAK_GINST, "unit": "Ginst/s",
                       "frac": round(rate / VALU_PEAK_GINST, 4), "insts_per_launch": valu}
        if out["valu"]["frac"] > out["frac"]:
            out.update(bound="valu", achieved=out["valu"]["achieved"], peak=VALU_PEAK_GINST, unit="Ginst/s",
                       frac=out["valu"]["frac"])
    if sweeps > 1:
        out["effective_bw"] = {"value": round(alg_bytes / t / 1e9, 2), "unit": "GB/s",
                               "meaning": f"{sweeps} sweeps' algorithmic bytes per launch / launch time; "
                                          "above HBM because the intermediate sweeps stay on chip"}
    if c.get("summary"):
        out["profile"] = "profiles/" + c["summary"]
    if out["frac"] > 1.0:
        # a working set held in the 256 MB Infinity Cache can be served
        # above the HBM peak; say so instead of reporting a fraction > 1
        out["error"] = "frac > 1: the launch is served from the Infinity Cache or its counter figures are stale"
        log(f"roofline of {name}@{shape}: {out['error']}")
    return out


def cpu_threads() -> int:
    """Host threads for the OpenMP baseline: the CPU share of this process
    (OMP_NUM_THREADS when the environment sets one, else every core in the
    affinity mask)."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, len(os.sched_getaffinity(0)))


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_rates(run, units0, budget_s, max_units, unit_lups):
    """Times `run(threads, units)` (the oracle over `units` steps) serially
    and on cpu_threads() threads, each sized to about `budget_s` seconds
    from a probe of `units0` steps.  Returns {threads: (GLUPS, units, s)}."""
    out = {}
    for th in (1, cpu_threads()):
        if th in out:
            continue
        t0 = time.perf_counter()
        run(th, units0)
        dt = time.perf_counter() - t0
        units = max(units0, min(max_units, int(units0 * budget_s / max(dt, 1e-3))))
        t0 = time.perf_counter()
        run(th, units)
        dt = time.perf_counter() - t0
        out[th] = (1e-9 * unit_lups * units / dt, units, dt)
    return out


def cpu_entry(rates, what):
    """The cpu_baseline object: the OpenMP rate as value, the serial one
    beside it, host facts stated."""
    th = max(rates)
    v, units, dt = rates[th]
    sv, sunits, sdt = rates[1]
    return {"value": round(v, 4), "unit": "GLUPS", "cores": th, "kind": "port",
            "sample": f"{what}: OpenMP {th} threads {units} steps in {dt:.1f}s; serial {sunits} steps in {sdt:.1f}s",
            "serial": {"value": round(sv, 4), "unit": "GLUPS", "cores": 1},
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}


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

    def __init__(self, n=16384, generations=1000, seed=synthetic.DEFAULT_SEED, per_launch=31, variant=7, wrap=0):
        self.n, self.gens, self.seed, self.wrap = n, generations, seed, wrap
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

    def topology(self):
        return {0: "dead Cube edge", 1: "x wrapped, dead edge in y", 2: "y wrapped, dead edge in x",
                3: "Torus"}[self.wrap]

    def config(self):
        c = {"workload": f"Conway Game of Life 2D uint8 {self.n}x{self.n}, Bernoulli(0.3), {self.topology()}, "
                         f"{self.gens} generations", "grid": [self.n, self.n], "generations": self.gens,
             "seed": self.seed, "generations_per_launch": self.per_launch, "parallelism": "single GPU"}
        if self.wrap:
            c["wrap_mask"] = self.wrap
        return c

    def lups_per_step(self):
        return float(self.n) * self.n * self.gens

    def step(self):
        s = self.stream.h
        lgdhip.call("lgdhip_memcpy_d2d", self.work.ptr, self.pristine.ptr, self.nbytes, s)
        if not self.wrap:
            lgdhip.call("lgdhip_life2d_run", self.work.ptr, C.byref(self.g), 0, self.gens, self.ws.ptr,
                        self.ws_bytes, C.byref(self.tile), s)
            return
        # periodic: the same pack -> generations -> unpack as lgdhip_life2d_run
        pa, pb = self.ws.value, self.ws.value + self.ws_bytes // 2
        in_b = C.c_int()
        lgdhip.call("lgdhip_life2d_pack", self.work.ptr, C.byref(self.g), C.c_void_p(pa), s)
        lgdhip.call("lgdhip_life2d_packed_run_periodic", C.c_void_p(pa), C.c_void_p(pb), C.byref(self.g), 0,
                    self.gens, self.wrap, C.byref(self.tile), C.byref(in_b), s)
        lgdhip.call("lgdhip_life2d_unpack", C.c_void_p(pb if in_b.value else pa), C.byref(self.g), self.work.ptr, s)

    def dominant_kernel(self):
        """The roofline of the generation kernel, timed over a block of
        launches of `per_launch` generations each on the packed state."""
        s = self.stream.h
        half = self.ws_bytes // 2
        pa, pb = self.ws.value, self.ws.value + half
        lgdhip.call("lgdhip_life2d_pack", self.pristine.ptr, C.byref(self.g), C.c_void_p(pa), s)
        launches = 64
        gens = launches * self.per_launch
        in_b = C.c_int()
        if self.wrap:
            ms = timed(self.stream, lambda: lgdhip.call("lgdhip_life2d_packed_run_periodic", C.c_void_p(pa),
                                                        C.c_void_p(pb), C.byref(self.g), 0, gens, self.wrap,
                                                        C.byref(self.tile), C.byref(in_b), s))
        else:
            ms = timed(self.stream, lambda: lgdhip.call("lgdhip_life2d_packed_run", C.c_void_p(pa), C.c_void_p(pb),
                                                        C.byref(self.g), 0, gens, C.byref(self.tile),
                                                        C.byref(in_b), s))
        per_launch_bytes = self.bytes_per_lup * self.n * self.n * self.per_launch
        # variant 0 is the library default (a stacked-wave kernel)
        name = "life_packed_tb" if 1 <= self.tile.variant <= 4 else "life_packed_stacked"
        # one launch reads and writes the bit-packed state once (1 bit per cell)
        packed = 2.0 * self.n * self.n / 8
        return roofline(name, f"{self.n}x{self.n},{self.per_launch}gen", ms / launches, per_launch_bytes,
                        self.per_launch, packed)

    def cpu_baseline(self, budget_s=10.0):
        oracle = load_oracle()
        sample = self.host[:4096, :4096].copy()
        rates = cpu_rates(lambda th, k: oracle.life2d(sample, k, 0, th, wrap=self.wrap), 4, budget_s, 4000,
                          sample.size)
        return cpu_entry(rates, "C oracle (byte-per-cell rule) on the top-left 4096x4096 of the same input")


# ----------------------------------------------------------------------
# Jacobi3D FP64
# ----------------------------------------------------------------------
class Jacobi:
    name = "jacobi"
    dtype = "f64"
    bytes_per_lup = 16.0

    def __init__(self, n=128, steps=100, seed=synthetic.DEFAULT_SEED, variant=0, by=0, bz=0, flags=0, wrap=0, **_):
        self.n, self.sweeps, self.seed, self.wrap = n, steps, seed, wrap
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
        topo = "zero Cube edge" if not self.wrap else ("Torus" if self.wrap == 7 else f"wrap mask {self.wrap}, zero edge")
        c = {"workload": f"Jacobi3D 7-point FP64 {self.n}^3, {topo}, U[0,1) interior, {self.sweeps} steps",
             "grid": [self.n] * 3, "steps": self.sweeps, "seed": self.seed, "parallelism": "single GPU"}
        if self.wrap:
            c["wrap_mask"] = self.wrap
        if getattr(self, "plan", None):
            c["launch"] = self.plan  # the library's plan: kernel variant and z planes per work-group
        return c

    def lups_per_step(self):
        return float(self.n) ** 3 * self.sweeps

    def run(self, sweeps, s):
        if self.wrap:
            lgdhip.call("lgdhip_jacobi3d_f64_run_periodic", self.a.ptr, self.b.ptr, C.byref(self.g), sweeps,
                        self.wrap, 0.0, C.byref(self.tile), None, s)
        else:
            lgdhip.call("lgdhip_jacobi3d_f64_run", self.a.ptr, self.b.ptr, C.byref(self.g), sweeps,
                        C.byref(self.tile), None, s)

    def step(self):
        s = self.stream.h
        lgdhip.call("lgdhip_memcpy_d2d", self.a.ptr, self.pristine.ptr, self.nbytes, s)
        self.run(self.sweeps, s)

    # sweeps per launch of the temporal-blocking variants (9, 10 and 12 hold
    # two rows per thread; 11 and 12 run the same pipeline for one sweep);
    # the library resolves variant 0 by its launch plan
    TB_DEPTH = {4: 2, 5: 3, 6: 4, 7: 2, 8: 3, 9: 4, 10: 4, 11: 1, 12: 1}

    def dominant_kernel(self):
        v, zc = C.c_int(0), C.c_int(0)
        lgdhip.call("lgdhip_jacobi3d_f64_plan", C.byref(self.g), C.byref(self.tile), C.byref(v), C.byref(zc))
        variant = v.value
        k = self.TB_DEPTH.get(variant, 1)
        sweeps = 240   # a multiple of every blocking depth: launches of one kernel only
        ms = timed(self.stream, lambda: self.run(sweeps, self.stream.h))
        if variant in self.TB_DEPTH:
            name = "jacobi3d_f64_tb2" if variant in (9, 10, 12) else "jacobi3d_f64_tb"
        else:
            name = "jacobi3d_f64_zmarch" if self.tile.variant == 1 else "jacobi3d_f64_zbatch"
        self.plan = {"variant": variant, "zchunk": zc.value}
        # one launch advances k sweeps: algorithmic bytes are k sweeps' worth
        shape = f"{self.n}^3,K{k}" if k > 1 else f"{self.n}^3"
        return roofline(name, shape, ms / (sweeps // k), k * self.bytes_per_lup * float(self.n) ** 3, k)

    def cpu_baseline(self, budget_s=10.0):
        oracle = load_oracle()
        probe = max(1, int(16 * (128.0 / self.n) ** 3))  # about 1e-1 s serial
        rates = cpu_rates(lambda th, k: oracle.jacobi3d(self.host, k, 0.0, th, wrap=self.wrap), probe, budget_s,
                          200000, self.host.size)
        return cpu_entry(rates, f"C oracle, same {self.n}^3 input")


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
        shape = f"{self.n}^3" + (",zslabs" if self.g.layout == lgdhip.LAYOUT_Z_SLABS else "") + \
            (",2steps" if per == 2 else "")
        name = "lbm_d3q19_pull2" if per == 2 else "lbm_d3q19_pull"
        return roofline(name, shape, ms / launches, per * self.bytes_per_lup * float(self.n) ** 3, per)

    def cpu_baseline(self, budget_s=10.0):
        oracle = load_oracle()
        m = 64
        ux, uy, uz = synthetic.lbm_initial_velocity(m, m, m, 1e-3, self.seed)
        f0 = oracle.lbm_equilibrium(np.ones_like(ux), ux, uy, uz)
        fl = oracle.lbm_cavity_flags(m, m, m)
        rates = cpu_rates(lambda th, k: oracle.lbm_run(f0, fl, k, 1.7, (0.05, 0.0, 0.0), th), 2, budget_s, 2000,
                          m ** 3)
        return cpu_entry(rates, "C oracle FP64 D3Q19, 64^3 cavity")


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


def poisson_sweep_kernel(stream, e0, e1, f, probs, count, ws, ws_bytes, cells, shape, variant=0):
    """The roofline of the Poisson solver's sweep kernel: temporal blocking,
    4 sweeps per launch (12 B per sweep-LUP, so 4 x 12 B x cells per launch);
    variant 0 is the library default (6, three rows per thread, on large
    launches, else 3, one row).  Returns (roofline, planned variant)."""
    k, sweeps = 4, 48
    in_b = C.c_int(0)
    tile = lgdhip.tile(variant=variant)
    ms = timed(stream, lambda: lgdhip.call("lgdhip_poisson_sweeps_f32", e0.ptr, e1.ptr, f.ptr, probs, count, sweeps,
                                           C.byref(tile), ws.ptr, ws_bytes, C.byref(in_b), stream.h))
    planned = C.c_int(0)  # the library resolves the default by launch size
    lgdhip.call("lgdhip_poisson_sweep_plan", probs, count, variant, C.byref(planned))
    name = "poisson_sweep_tb2" if planned.value in (4, 5, 6) else "poisson_sweep_tb"
    return roofline(name, shape, ms / (sweeps // k), k * 12.0 * cells, k), planned.value


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
        roof, self.sweep_plan = poisson_sweep_kernel(self.stream, e0, e1, self.f, self.probs, self.count, self.ws,
                                                     self.ws_bytes, self.cells, f"{self.count}problems",
                                                     self.variant)
        return roof

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
                "sample": f"C oracle plain FP64 Jacobi (1 thread; the oracle's Poisson sweep is serial), one {m}^3 "
                          f"problem of the ensemble, {sweeps} sweeps, {dt:.1f}s", "cpu_model": cpu_model(),
                "host_cpus": os.cpu_count()}


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

    def __init__(self, n=128, steps=30000, seed=synthetic.DEFAULT_SEED, train=512, eval=16, lr=5e-4, weights="",
                 **_):
        from libgeodecomp_amd import learned_guess as lgm
        self.lgm = lgm
        self.n, self.train_steps, self.seed, self.ntrain, self.neval = n, steps, seed, train, eval
        self.lr = lr
        self.stream = lgdhip.Stream()
        self.weights = weights
        self.result = None
        if weights:
            # a network this library trained earlier (learned_guess --save):
            # no generation, no training; the timed step is the same
            self.model, meta = lgm.LearnedGuess.load(weights, stream=self.stream.h)
            self.weights_meta = meta
            self.n = self.model.n
            self.m = meta.get("coarse_m", lgm.default_coarse_size(self.n))
            self.ntrain, self.train_steps = meta.get("train_problems"), meta.get("train_steps")
            self.lr = meta.get("peak_lr")
            self.generate_s, self.train_s = meta.get("generate_s"), meta.get("train_s")
            self.log = [(self.train_steps, meta.get("final_loss"))]
            self.eval = lgm.Ensemble(synthetic.poisson_ensemble(eval, sizes=(self.n,), seed=seed + 7919),
                                     stream=self.stream.h)
            return
        self.m = lgm.default_coarse_size(n)
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

    def config(self):
        r = self.result or {}
        c = {"workload": f"learned initial guess vs zero guess, {self.neval} held-out Poisson problems {self.n}^3 "
                            f"(Gaussian sources, Dirichlet faces), Jacobi polishing to 1e-6 relative FP64 residual",
                "network": "3D U-Net 16-32-64-128, 3x3x3 BF16 convs on MFMA, FP32 accumulation, bias-free",
                "coarse_grid": self.m, "train_problems": self.ntrain, "train_steps": self.train_steps,
                "peak_lr": self.lr,
                "generate_s": self.generate_s and round(self.generate_s, 1),
                "train_s": self.train_s and round(self.train_s, 1),
                "final_loss": self.log[-1][1] if self.log else None, "arms": r.get("arms"),
                "sweep_ratio": r.get("sweep_ratio"), "fp64_polish_sweeps": r.get("fp64_polish_sweeps"),
                "network_gain_over_coarse_base": self.network_gain(), "seed": self.seed,
                "parallelism": "single GPU"}
        if self.weights:
            c["weights"] = {"file": os.path.relpath(self.weights, REPO), "trained_by": self.weights_meta.get("command")}
        return c

    def network_gain(self):
        """The network's own share: wall time of the coarse-base guess alone
        over the learned guess (both polished by the same solver)."""
        arms = {a["arm"]: a for a in (self.result or {}).get("arms", [])}
        if "coarse" not in arms or "learned" not in arms:
            return None
        return round(arms["coarse"]["wall_s"] / arms["learned"]["wall_s"], 3)

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
        """The polishing sweep over the held-out set (the blocked sweep)."""
        e0 = lgdhip.DeviceBuffer(self.eval.elems * 4)
        e1 = lgdhip.DeviceBuffer(self.eval.elems * 4)
        for b in (e0, e1):
            lgdhip.call("lgdhip_fill_u8", b.ptr, 0, self.eval.elems * 4, self.stream.h)
        roof, self.sweep_plan = poisson_sweep_kernel(self.stream, e0, e1, self.eval.f, self.eval.probs,
                                                     self.eval.count, self.eval.ws, self.eval.ws_bytes,
                                                     self.neval * float(self.n) ** 3, f"{self.neval}x{self.n}^3")
        return roof

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
                          f"{dt:.1f}s (sweep rate; the zero-guess solve needs ~3e4 sweeps per problem)",
                "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}


class UNet:
    """The learned guess's network alone: 3D U-Net 16-32-64-128 on one 128^3
    sample (3 input channels: f, b and the coarse base), BF16 operands on
    the MFMA cores with FP32 accumulation, random weights and inputs.  One
    step = one training step on one sample (forward, MSE loss, backward,
    Adam); the forward alone is timed beside it.  FLOPs count the
    convolutions only (2 x 27 x cin x cout per voxel; backward = 2 x
    forward)."""
    name = "unet"
    dtype = "bf16"

    def __init__(self, n=128, seed=synthetic.DEFAULT_SEED, **_):
        from libgeodecomp_amd.unet import UNet3D
        self.n = n
        self.stream = lgdhip.Stream()
        s = self.stream.h
        self.net = UNet3D(n=n, in_ch=3, seed=seed % 1000, stream=s, bias=False)
        rng = np.random.default_rng(seed)
        self.fields = []
        for _ in range(3):
            b = lgdhip.DeviceBuffer(n ** 3 * 4)
            b.upload(rng.standard_normal(n ** 3).astype(np.float32), s)
            self.fields.append(b)
        self.target = lgdhip.DeviceBuffer(n ** 3 * 4)
        self.target.upload(rng.standard_normal(n ** 3).astype(np.float32), s)
        self.net.load_input(self.fields)
        self.fwd_flops = self.net.flops_forward()
        self.fwd_ms = None

    def config(self):
        return {"workload": f"3D U-Net 16-32-64-128 training step on one {self.n}^3 sample (3 input channels), "
                            "BF16 MFMA convolutions, FP32 accumulation", "grid": [self.n] * 3,
                "forward_gflop": round(self.fwd_flops / 1e9, 1), "forward_ms": self.fwd_ms,
                "forward_tflops": round(self.fwd_flops / self.fwd_ms * 1e-9, 1) if self.fwd_ms else None,
                "parallelism": "single GPU"}

    def step(self):
        self.net.forward()
        self.net.loss(self.target, read=False)
        self.net.backward()
        self.net.adam(lr=1e-6)

    def value(self):
        """(TFLOP/s of the training step, unit, ms per step) -- set by
        measure() through lups_per_step()."""
        return self._value

    def lups_per_step(self):
        return 0.0

    def dominant_kernel(self):
        """The forward pass as one unit (its 15 convolutions, 3 poolings, 3
        upsamplings): BF16 MFMA roofline over the convolution FLOPs."""
        reps = 10
        ms = timed(self.stream, self.net.forward, reps=reps) / reps
        self.fwd_ms = round(ms, 4)
        achieved = self.fwd_flops / (ms * 1e-3) / 1e12
        bound_us, mfma_us = self.layer_bound_us()
        return {"kernel": "unet3d_forward (15 convolutions: z-marching, LDS-tile and GEMM-tiled MFMA kernels)",
                "bound": "mfma",
                "achieved": round(achieved, 2), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / MFMA_BF16_PEAK_TFLOPS, 4), "traffic": None,
                "avg_launch_us": round(ms * 1e3, 3), "algorithmic_flops_per_launch": self.fwd_flops,
                "layers": {"bound_us": round(bound_us, 2), "mfma_only_us": round(mfma_us, 2),
                           "frac": round(bound_us / (ms * 1e3), 4),
                           "meaning": "sum over the forward's 15 convolutions, 3 poolings and 3 upsamplings of "
                                      "max(FLOPs / 2.5 PFLOP/s, bytes each must move at least once / 8 TB/s), "
                                      "over the measured forward time: most 128^3 and 64^3 layers are "
                                      "HBM-bound at these channel counts"}}

    def layer_bound_us(self):
        """The forward's per-layer roofline: each op costs at least the larger
        of its FLOPs at the MFMA peak and its compulsory bytes (BF16 inputs
        read once, outputs written once; the last layer writes FP32) at the
        HBM peak.  Returns (that sum, the MFMA-only sum) in us."""
        net = self.net
        c0, c1, c2, c3 = net.ch
        n = self.n
        total = mfma = 0.0

        def op(flops, nbytes):
            nonlocal total, mfma
            t_mfma = flops / (MFMA_BF16_PEAK_TFLOPS * 1e12) * 1e6
            total += max(t_mfma, nbytes / (HBM_PEAK_GBS * 1e9) * 1e6)
            mfma += t_mfma

        for name in net.order:
            l = net.layers[name]
            taps = 27 if l.ksize == 3 else 1
            out_bytes = 4 if name == "out" else 2
            op(2.0 * taps * l.cin * l.cout * l.n ** 3, l.n ** 3 * (l.cin * 2 + l.cout * out_bytes))
        for m, c in ((n, c0), (n // 2, c1), (n // 4, c2)):   # poolings
            op(0.0, m ** 3 * c * 2 + (m // 2) ** 3 * c * 2)
        for m, c in ((n // 8, c3), (n // 4, c2), (n // 2, c1)):   # upsamplings
            op(0.0, m ** 3 * c * 2 + (2 * m) ** 3 * c * 2)
        return total, mfma

    def cpu_baseline(self, budget_s=0.0):
        return None


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
        return roofline("lgd_copy16_kernel", f"{self.n}B", ms / launches, 2.0 * self.n)

    def cpu_baseline(self, budget_s=0.0):
        return None


WORKLOADS = {"life": Life, "jacobi": Jacobi, "lbm": LBM, "poisson": Poisson, "learned": Learned, "copy": Copy,
             "unet": UNet}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default=None,
                    help="one workload (default: life, plus the secondary Jacobi and LBM entries)")
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
    ap.add_argument("--wrap", type=int, default=0,
                    help="jacobi (bits 1, 2, 4) / life (bits 1, 2): periodic axes (0: the Cube)")
    ap.add_argument("--weights", default="", help="learned: evaluate a saved network instead of training one")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="default run: the main workload only")
    args = ap.parse_args()
    default_run = args.workload is None and not (args.size or args.iters or args.variant >= 0 or args.wrap)
    if args.workload is None:
        args.workload = "life"

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
        kw.update(per_launch=args.per_launch, variant=7 if args.variant < 0 else args.variant, wrap=args.wrap)
    flags = (lgdhip.TILE_NO_XCD_SWIZZLE if args.no_xcd_swizzle else 0) | \
        (lgdhip.TILE_XCD_SWIZZLE if args.xcd_swizzle else 0) | (lgdhip.TILE_XCD_ROWS if args.xcd_rows else 0)
    if args.workload == "jacobi":
        kw.update(variant=max(0, args.variant), by=args.by, bz=args.bz, flags=flags, wrap=args.wrap)
    if args.workload == "lbm":
        kw.update(by=args.by, flags=flags, variant=max(0, args.variant), bx=args.bx, bz=args.bz)
        if args.layout:
            kw.update(layout=lgdhip.LAYOUT_Z_SLABS if args.layout == "zslabs" else lgdhip.LAYOUT_MEMBER_PLANES)
    if args.workload == "poisson":
        kw.update(use_graph=0 if args.no_graph else 1, tol=args.tol, variant=max(0, args.variant))
    if args.workload == "learned" and args.weights:
        w = Path(args.weights)
        kw.update(weights=str(w if w.is_absolute() or w.exists() else REPO / w))  # else relative to the repository
    wl = WORKLOADS[args.workload](**kw)
    main_entry = measure(wl, args.steps, args.warmup, not args.no_cpu_baseline, 10.0)
    secondary = []
    if default_run and not args.no_secondary:
        # the other sweep configurations, each timed the same way on its own
        # shape, so the line carries their GLUPS and rooflines too
        for sub, skw in SECONDARY:
            w2 = WORKLOADS[sub](**skw)
            e = measure(w2, max(args.steps, 2), max(args.warmup, 1), not args.no_cpu_baseline, 5.0)
            secondary.append({"workload": e["config"]["workload"], "value": e["value"], "unit": e["unit"],
                              "ms_per_step": e["ms_per_step"], "steps": e["steps"], "dtype": w2.dtype,
                              "config": e["config"], "roofline": e["roofline"], "cpu_baseline": e["cpu_baseline"]})
            del w2
            lgdhip.call("lgdhip_device_synchronize")

    line = {
        "metric": METRIC,
        "value": main_entry["value"],
        "unit": main_entry["unit"],
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_entry["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": wl.dtype,
        "data": "synthetic (SplitMix64 seeded initializer)",
        "config": main_entry["config"],
        "roofline": main_entry["roofline"],
        "cpu_baseline": main_entry["cpu_baseline"],
    }
    if secondary:
        line["secondary"] = secondary
    print(json.dumps(line), flush=True)


# the secondary entries of the default run: Jacobi3D FP64 128^3 x 100 (the
# configured shape, Infinity-Cache resident), Jacobi3D 512^3 x 100 (the
# HBM-bound shape), D3Q19 LBM FP32 256^3 x 2000
SECONDARY = [("jacobi", dict(n=128, steps=100)), ("jacobi", dict(n=512, steps=100)),
             ("lbm", dict(n=256, steps=2000)), ("unet", dict(n=128))]
# the learned initial guess against the zero guess on 16 held-out 128^3
# problems, with the network this library trained and saved (training takes
# minutes; the command that made the file is in the entry's config)
LEARNED_WEIGHTS = REPO / "libgeodecomp_amd" / "weights" / "unet_guess_128.npz"
if LEARNED_WEIGHTS.exists():
    SECONDARY.append(("learned", dict(weights=str(LEARNED_WEIGHTS))))


def measure(wl, steps, warmup, cpu, cpu_budget):
    """Warm-up, `steps` timed steps, the dominant kernel's roofline and the
    CPU baseline of one workload."""
    for _ in range(warmup):
        wl.step()
    wl.stream.sync()
    log(f"{wl.name}: warmup {warmup} done")
    lgdhip.call("lgdhip_device_synchronize")
    t0 = time.perf_counter()
    ms = timed(wl.stream, wl.step, reps=steps)
    lgdhip.call("lgdhip_device_synchronize")
    wall = time.perf_counter() - t0
    ms_per_step = ms / steps
    unit = "GLUPS"
    if isinstance(wl, UNet):
        wl._value = (3.0 * wl.fwd_flops / (ms_per_step * 1e-3) / 1e12, "TFLOP/s (training step)", ms_per_step)
    if hasattr(wl, "value"):
        value, unit, ms_per_step = wl.value()
    else:
        value = 1e-9 * wl.lups_per_step() / (ms_per_step * 1e-3)
    log(f"{wl.name}: timed {steps} steps: {ms:.3f} ms (wall {wall * 1e3:.1f} ms), {value:.3f} {unit}")
    roof = wl.dominant_kernel()
    log(f"{wl.name}: dominant kernel {roof['kernel']}: {roof['avg_launch_us']:.2f} us/launch, bound {roof['bound']}, "
        f"{roof['achieved']} {roof['unit']}, frac {roof['frac']}")
    base = None
    if cpu:
        base = wl.cpu_baseline(cpu_budget)
        log(f"{wl.name}: cpu baseline: {base}")
    return {"value": round(value, 3), "unit": unit, "ms_per_step": round(ms_per_step, 4), "steps": steps,
            "config": wl.config(), "roofline": roof, "cpu_baseline": base}


if __name__ == "__main__":
    main()
