"""Batched Poisson ensemble: oracle pinning (CPU) and GPU parity.

The reference has no Poisson solver; the oracle is pinned by the closed-form
Jacobi spectrum, manufactured integer solutions and a brute-force loop.
"""
import ctypes as C

import numpy as np
import pytest

from libgeodecomp_amd import synthetic


# ----------------------------------------------------------------------
# helpers
# ----------------------------------------------------------------------
def brute_sweep_f32(e, r):
    n = e.shape[0] - 2
    out = np.zeros_like(e)
    for z in range(1, n + 1):
        for y in range(1, n + 1):
            for x in range(1, n + 1):
                v = np.float32(e[z - 1, y, x])
                v = np.float32(v + e[z, y - 1, x])
                v = np.float32(v + e[z, y, x - 1])
                v = np.float32(v + e[z, y, x + 1])
                v = np.float32(v + e[z, y + 1, x])
                v = np.float32(v + e[z + 1, y, x])
                v = np.float32(v + r[z, y, x])
                out[z, y, x] = np.float32(v * np.float32(1.0 / 6.0))
    return out


def apply_A(u):
    """A u = 6u - sum_nb u over the interior of a padded block (FP64)."""
    c = u[1:-1, 1:-1, 1:-1]
    nb = (u[:-2, 1:-1, 1:-1] + u[2:, 1:-1, 1:-1] + u[1:-1, :-2, 1:-1] + u[1:-1, 2:, 1:-1]
          + u[1:-1, 1:-1, :-2] + u[1:-1, 1:-1, 2:])
    return 6.0 * c - nb


def padded(n, faces, oracle):
    u = np.zeros((n + 2,) * 3)
    oracle.poisson_faces(u, faces)
    return u


# ----------------------------------------------------------------------
# oracle pinning (CPU)
# ----------------------------------------------------------------------
def test_oracle_sweep_matches_brute_force(oracle):
    rng = np.random.default_rng(1)
    e = np.zeros((6, 6, 6), np.float32)
    e[1:-1, 1:-1, 1:-1] = rng.standard_normal((4, 4, 4))
    r = np.zeros_like(e)
    r[1:-1, 1:-1, 1:-1] = rng.standard_normal((4, 4, 4))
    assert np.array_equal(oracle.poisson_sweep_f32(e, r), brute_sweep_f32(e, r))


def test_oracle_damped_sweep_pins(oracle):
    """The weighted sweep e' = fma(w, fma(sum, 1/6, -e), e): w = 0 returns e
    bit for bit; any w agrees with e + w (J(e) - e) evaluated in FP64 from the
    oracle's own plain sweep to a few FP32 ulps; the halo stays untouched."""
    rng = np.random.default_rng(21)
    n = 9
    e = np.zeros((n + 2,) * 3, np.float32)
    e[1:-1, 1:-1, 1:-1] = rng.standard_normal((n, n, n))
    r = rng.standard_normal((n + 2,) * 3).astype(np.float32)
    assert np.array_equal(oracle.poisson_sweep_damped_f32(e, r, 0.0), e)
    plain = oracle.poisson_sweep_f32(e, r).astype(np.float64)
    for w in (2.0 / 3.0, 0.5, 1.0):
        got = oracle.poisson_sweep_damped_f32(e, r, w).astype(np.float64)
        exp = e.astype(np.float64) + np.float64(np.float32(w)) * (plain - e)
        inner = (slice(1, -1),) * 3
        assert np.all(np.abs(got[inner] - exp[inner]) <= 4 * 2.0 ** -23 * (np.abs(exp[inner]) + np.abs(e[inner]) + 1e-30))
        halo = np.ones(e.shape, bool)
        halo[inner] = False
        assert np.all(got[halo] == 0)


def test_oracle_jacobi_decay_rate_closed_form(oracle):
    """For the smoothest eigenvector v of A (product of sines, zero faces),
    f = A v and u0 = 0, the residual after k sweeps is rho^k r0 with
    rho = cos(pi/(n+1)): the Jacobi spectrum in closed form."""
    n = 10
    i = np.arange(1, n + 1)
    s = np.sin(np.pi * i / (n + 1))
    v = np.zeros((n + 2,) * 3)
    v[1:-1, 1:-1, 1:-1] = s[:, None, None] * s[None, :, None] * s[None, None, :]
    # f = A v exactly representable?  compute in FP64, store FP32 then use
    # that f: the rate argument holds for any f with f = A v_f32
    f = np.zeros((n + 2,) * 3, np.float32)
    f[1:-1, 1:-1, 1:-1] = apply_A(v).astype(np.float32)
    u = np.zeros((n + 2,) * 3)
    r0 = oracle.poisson_residual_norm(u, f)
    rho = np.cos(np.pi / (n + 1))
    for k in (1, 5, 20):
        uk = np.zeros((n + 2,) * 3)
        sweeps, rk = oracle.poisson_jacobi(uk, f, rel_tol=1e-300, max_sweeps=k, check=k)
        assert sweeps == k
        # FP32 rounding of f excites other modes at the 1e-8 level
        assert abs(rk / r0 - rho ** k) < 1e-6


def test_oracle_recovers_manufactured_integer_solution(oracle):
    rng = np.random.default_rng(7)
    n = 9
    faces = rng.integers(-3, 4, size=6).astype(np.float64)
    ustar = padded(n, faces, oracle)
    ustar[1:-1, 1:-1, 1:-1] = rng.integers(-20, 21, size=(n, n, n))
    f = np.zeros((n + 2,) * 3, np.float32)
    f[1:-1, 1:-1, 1:-1] = apply_A(ustar)          # integers: exact in FP32
    assert oracle.poisson_residual_norm(ustar, f) == 0.0
    u = padded(n, faces, oracle)
    sweeps, rn = oracle.poisson_jacobi(u, f, rel_tol=1e-10)
    r0 = oracle.poisson_residual_norm(padded(n, faces, oracle), f)
    assert rn <= 1e-10 * r0
    # error bound: |e|_2 <= |A^-1|_2 |r|_2 with lambda_min = 6 - 6 cos(pi/(n+1))
    lam = 6.0 - 6.0 * np.cos(np.pi / (n + 1))
    assert np.linalg.norm(u - ustar) <= rn / lam * 1.0001 + 1e-12


# ----------------------------------------------------------------------
# GPU
# ----------------------------------------------------------------------
class Batch:
    """Pools and workspace of one ensemble on the device."""

    def __init__(self, hip, problems, guesses=None):
        self.hip = hip
        self.count = len(problems)
        self.probs = (hip.lgdhip_poisson_problem * self.count)()
        for k, (n, faces, _) in enumerate(problems):
            self.probs[k].n = n
            for j in range(6):
                self.probs[k].face[j] = faces[j]
        self.elems = hip.load().lgdhip_poisson_layout(self.probs, self.count)
        self.ws_bytes = hip.load().lgdhip_poisson_workspace_bytes(self.probs, self.count)
        self.ws = hip.DeviceBuffer(self.ws_bytes)
        self.u = hip.DeviceBuffer(self.elems * 8)
        self.f = hip.DeviceBuffer(self.elems * 4)
        u_host = np.zeros(self.elems)
        f_host = np.zeros(self.elems, np.float32)
        for k, (n, faces, f) in enumerate(problems):
            o, m = self.probs[k].offset, (n + 2) ** 3
            f_host[o:o + m] = f.ravel()
            if guesses is not None:
                u_host[o:o + m] = guesses[k].ravel()
        self.u.upload(u_host)
        self.f.upload(f_host)
        hip.call("lgdhip_poisson_set_faces", self.u.ptr, self.probs, self.count, self.ws.ptr, self.ws_bytes, None)

    def solve(self, rel_tol=1e-6, chunk=32, use_graph=1, reference=None, stream=None, sweep_variant=0, polish=0):
        hip = self.hip
        opt = hip.lgdhip_poisson_options()
        opt.rel_tol, opt.chunk, opt.max_sweeps, opt.use_graph = rel_tol, chunk, 2_000_000, use_graph
        opt.sweep_variant = sweep_variant
        opt.final_fp64_sweeps = polish
        ref = None
        if reference is not None:
            ref = (C.c_double * self.count)(*reference)
            opt.reference_norms = ref
        res = (hip.lgdhip_poisson_result * self.count)()
        hip.call("lgdhip_poisson_batch_solve", self.u.ptr, self.f.ptr, self.probs, self.count, C.byref(opt),
                 self.ws.ptr, self.ws_bytes, res, stream)
        return res

    def u_blocks(self):
        host = np.empty(self.elems)
        self.u.download(host)
        out = []
        for k in range(self.count):
            n, o = self.probs[k].n, self.probs[k].offset
            out.append(host[o:o + (n + 2) ** 3].reshape((n + 2,) * 3).copy())
        return out


@pytest.mark.gpu
@pytest.mark.parametrize("variant,sweeps", [(1, 5), (2, 5), (3, 9), (3, 8), (0, 4), (4, 9), (4, 8), (5, 12),
                                            (6, 9), (6, 16)])
def test_gpu_sweeps_bit_exact(hip, oracle, variant, sweeps):
    """Plain sweeps (variant 1) and temporal blocking with 2 (variant 2) or 4
    (variant 3) sweeps per launch, 4 with two rows per thread (variants 4
    and 5: 32- and 16-row tiles) or three (variant 6, the default for large
    launches: 48-row tiles), sweep counts with remainders,
    sizes from 1 to several tiles per axis, noise in r's halo: bit-identical
    to the oracle."""
    rng = np.random.default_rng(3 + variant)
    sizes = [1, 5, 8, 17, 32, 33, 40, 70, 113]
    problems = [(n, np.zeros(6), np.zeros((n + 2,) * 3, np.float32)) for n in sizes]
    b = Batch(hip, problems)
    e_host = np.zeros(b.elems, np.float32)
    r_host = np.zeros(b.elems, np.float32)
    blocks = []
    for k, n in enumerate(sizes):
        o = b.probs[k].offset
        e = np.zeros((n + 2,) * 3, np.float32)
        # r's halo holds noise: it may only ever reach cells outside the problem
        r = (100 * rng.standard_normal((n + 2,) * 3)).astype(np.float32)
        e[1:-1, 1:-1, 1:-1] = rng.standard_normal((n, n, n))
        r[1:-1, 1:-1, 1:-1] = rng.standard_normal((n, n, n))
        e_host[o:o + e.size] = e.ravel()
        r_host[o:o + r.size] = r.ravel()
        blocks.append((e, r))
    ea, eb, r = (hip.DeviceBuffer(b.elems * 4) for _ in range(3))
    ea.upload(e_host)
    eb.upload(np.zeros(b.elems, np.float32))
    r.upload(r_host)
    in_b = C.c_int(-1)
    print("[case] poisson sweeps", sizes, variant, sweeps, flush=True)
    tile = hip.tile(variant=variant)
    hip.call("lgdhip_poisson_sweeps_f32", ea.ptr, eb.ptr, r.ptr, b.probs, b.count, sweeps, C.byref(tile), b.ws.ptr,
             b.ws_bytes, C.byref(in_b), None)
    out = np.empty(b.elems, np.float32)
    (eb if in_b.value else ea).download(out)
    for k, (e, rr) in enumerate(blocks):
        for _ in range(sweeps):
            e = oracle.poisson_sweep_f32(e, rr)
        o = b.probs[k].offset
        assert np.array_equal(out[o:o + e.size].reshape(e.shape), e), sizes[k]


@pytest.mark.gpu
@pytest.mark.parametrize("variant,sweeps", [(1, 3), (3, 9), (4, 10), (5, 8), (0, 12), (6, 9)])
def test_gpu_damped_sweeps_bit_exact(hip, oracle, variant, sweeps):
    """Weighted sweeps (omega = 2/3, the smoothing phase's) through every
    kernel -- single sweeps, blocked with one or two rows per thread, sweep
    counts with remainders -- equal the oracle's damped sweep bit for bit."""
    rng = np.random.default_rng(40 + variant)
    sizes = [1, 6, 17, 33, 70]
    problems = [(n, np.zeros(6), np.zeros((n + 2,) * 3, np.float32)) for n in sizes]
    b = Batch(hip, problems)
    e_host = np.zeros(b.elems, np.float32)
    r_host = np.zeros(b.elems, np.float32)
    blocks = []
    for k, n in enumerate(sizes):
        o = b.probs[k].offset
        e = np.zeros((n + 2,) * 3, np.float32)
        r = (10 * rng.standard_normal((n + 2,) * 3)).astype(np.float32)
        e[1:-1, 1:-1, 1:-1] = rng.standard_normal((n, n, n))
        e_host[o:o + e.size] = e.ravel()
        r_host[o:o + r.size] = r.ravel()
        blocks.append((e, r))
    ea, eb, r = (hip.DeviceBuffer(b.elems * 4) for _ in range(3))
    ea.upload(e_host)
    eb.upload(np.zeros(b.elems, np.float32))
    r.upload(r_host)
    in_b = C.c_int(-1)
    omega = 2.0 / 3.0
    print("[case] poisson damped sweeps", sizes, variant, sweeps, flush=True)
    tile = hip.tile(variant=variant)
    hip.call("lgdhip_poisson_damped_sweeps_f32", ea.ptr, eb.ptr, r.ptr, b.probs, b.count, sweeps, omega,
             C.byref(tile), b.ws.ptr, b.ws_bytes, C.byref(in_b), None)
    out = np.empty(b.elems, np.float32)
    (eb if in_b.value else ea).download(out)
    for k, (e, rr) in enumerate(blocks):
        for _ in range(sweeps):
            e = oracle.poisson_sweep_damped_f32(e, rr, omega)
        o = b.probs[k].offset
        assert np.array_equal(out[o:o + e.size].reshape(e.shape), e), sizes[k]


def check_solution(oracle, problem, u, result):
    """A solved problem (padded FP64 block u) meets the FP64 residual
    criterion as recomputed on the CPU and reports that residual; returns
    (zero-guess block u0, its residual norm r0, the solution's rk)."""
    n, faces, f = problem
    assert result.converged == 1
    u0 = padded(n, faces, oracle)
    r0 = oracle.poisson_residual_norm(u0, f)
    rk = oracle.poisson_residual_norm(u, f)
    assert rk <= 1e-6 * r0 * 1.0000001, (n, rk / r0)
    assert abs(result.residual_norm - rk) <= 1e-9 * r0
    return u0, r0, rk


@pytest.mark.gpu
@pytest.mark.parametrize("use_graph,variant", [(1, 0), (0, 0), (1, 1), (1, 2), (0, 3), (1, 4), (1, 5), (1, 6),
                                               (0, 6)])
def test_gpu_ensemble_solve(hip, oracle, use_graph, variant):
    """A small uneven ensemble (config-4 generator, sizes 8..40) through the
    solver with the default sweep kernel and each forced one (1 plain
    sweeps, 2 / 3 blocked one-row, 4 / 5 two rows, 6 three rows), graph and
    direct launches: every GPU solution meets the FP64 residual criterion as
    recomputed on the CPU, takes about as many sweeps as plain FP64 Jacobi,
    and agrees with the oracle's solution within the error bound the
    tolerance implies."""
    problems = synthetic.poisson_ensemble(10, sizes=(8, 16, 24, 32, 40), seed=11)
    print("[case] poisson ensemble", [p[0] for p in problems], "graph", use_graph, "variant", variant, flush=True)
    b = Batch(hip, problems)
    # graphs are captured on a non-null stream only
    stream = hip.Stream() if use_graph else None
    res = b.solve(rel_tol=1e-6, chunk=16, use_graph=use_graph, stream=stream.h if stream else None,
                  sweep_variant=variant)
    ublocks = b.u_blocks()
    for k, (n, faces, f) in enumerate(problems):
        ublock = ublocks[k]
        u0, r0, rk = check_solution(oracle, problems[k], ublock, res[k])
        uo = u0.copy()
        sweeps_o, rn_o = oracle.poisson_jacobi(uo, f, rel_tol=1e-6, check=1)
        # same linear iteration: the sweep count may exceed the oracle's by
        # the chunk granularity and the FP32 correction floor, not more
        assert sweeps_o <= res[k].sweeps <= 1.10 * sweeps_o + 2 * 16 * (res[k].outer_iterations + 1), \
            (k, n, sweeps_o, res[k].sweeps)
        lam = 6.0 - 6.0 * np.cos(np.pi / (n + 1))
        bound = (rk + rn_o) / lam
        assert np.linalg.norm(ublock - uo) <= 1.01 * bound + 1e-12
        print(f"  n={n}: sweeps gpu {res[k].sweeps} oracle {sweeps_o}, outer {res[k].outer_iterations}, "
              f"rel residual {rk / r0:.3e}", flush=True)


@pytest.mark.gpu
def test_gpu_manufactured_solution(hip, oracle):
    rng = np.random.default_rng(5)
    problems, exact = [], []
    for n in (7, 12, 30):
        faces = rng.integers(-3, 4, size=6).astype(np.float64)
        ustar = padded(n, faces, oracle)
        ustar[1:-1, 1:-1, 1:-1] = rng.integers(-20, 21, size=(n, n, n))
        f = np.zeros((n + 2,) * 3, np.float32)
        f[1:-1, 1:-1, 1:-1] = apply_A(ustar)
        problems.append((n, faces, f))
        exact.append(ustar)
    b = Batch(hip, problems)
    res = b.solve(rel_tol=1e-9, chunk=8)
    for k, u in enumerate(b.u_blocks()):
        n = problems[k][0]
        lam = 6.0 - 6.0 * np.cos(np.pi / (n + 1))
        assert res[k].converged == 1
        assert np.linalg.norm(u - exact[k]) <= res[k].residual_norm / lam * 1.01 + 1e-9


@pytest.mark.gpu
def test_gpu_guess_with_reference_norm(hip, oracle):
    """Starting from a good guess with the zero-guess reference norm needs
    fewer sweeps than from zero (the learned-guess use)."""
    problems = synthetic.poisson_ensemble(3, sizes=(24,), seed=2)
    zero = Batch(hip, problems)
    refs = np.zeros(3)
    hip.call("lgdhip_poisson_residual_norms", zero.u.ptr, zero.f.ptr, zero.probs, 3, zero.ws.ptr, zero.ws_bytes,
             refs.ctypes.data_as(C.POINTER(C.c_double)), None)
    res0 = zero.solve(rel_tol=1e-6, chunk=8)
    sol = zero.u_blocks()
    # a guess 1e-3 away from the solution in a smooth way
    guesses = []
    for k, (n, faces, f) in enumerate(problems):
        g = sol[k].copy()
        g[1:-1, 1:-1, 1:-1] *= 1.0 + 1e-3
        guesses.append(g)
    warm = Batch(hip, problems, guesses)
    res1 = warm.solve(rel_tol=1e-6, chunk=8, reference=refs)
    for k in range(3):
        assert res1[k].converged == 1 and res1[k].reference_norm == refs[k]
        assert res1[k].sweeps < res0[k].sweeps


@pytest.mark.gpu
def test_gpu_ensemble_with_largest_size_and_kernel_switch(hip, oracle):
    """2048 small problems (n = 8..16) and one of the configuration's
    largest size, n = 192: the active set starts at or above the three-row
    kernel's launch threshold (the plan query reports variant 6) and falls
    below it once the small problems have converged, so the solve re-plans
    to the one-row kernel (variant 3), re-uploads its tiles and re-captures
    its graph mid-solve.  Every problem meets the FP64 residual criterion
    recomputed on the CPU; a sample of the small ones also agrees with the
    oracle's plain FP64 Jacobi solution within the tolerance's bound (the
    oracle needs ~10^5 sweeps of 7 M cells for n = 192, so that one is held
    to the residual criterion alone)."""
    small = synthetic.poisson_ensemble(2048, sizes=(8, 12, 16), seed=71)
    big = synthetic.poisson_ensemble(1, sizes=(192,), seed=72)
    problems = small + big
    b = Batch(hip, problems)
    plan = C.c_int(0)
    hip.call("lgdhip_poisson_sweep_plan", b.probs, b.count, 0, C.byref(plan))
    assert plan.value == 6
    one = (hip.lgdhip_poisson_problem * 1)(b.probs[b.count - 1])
    hip.call("lgdhip_poisson_sweep_plan", one, 1, 0, C.byref(plan))
    assert plan.value == 3
    print("[case] poisson 2048 small + n=192, kernel switch", flush=True)
    stream = hip.Stream()
    res = b.solve(rel_tol=1e-6, chunk=64, use_graph=1, stream=stream.h)
    ublocks = b.u_blocks()
    for k in range(b.count):
        ublock = ublocks[k]
        u0, r0, rk = check_solution(oracle, problems[k], ublock, res[k])
        if k % 256 == 0 and k < len(small):
            n, _, f = problems[k]
            uo = u0.copy()
            sweeps_o, rn_o = oracle.poisson_jacobi(uo, f, rel_tol=1e-6, check=1)
            lam = 6.0 - 6.0 * np.cos(np.pi / (n + 1))
            assert np.linalg.norm(ublock - uo) <= 1.01 * (rk + rn_o) / lam + 1e-12
    big_sweeps = res[b.count - 1].sweeps
    # the smoothest mode decays by cos(pi/193) per sweep: ~1e5 sweeps for 1e-6
    print(f"  n=192: {big_sweeps} sweeps, {res[b.count - 1].outer_iterations} outer steps", flush=True)
    assert 3e4 < big_sweeps < 2e5
    assert max(r.sweeps for r in res[:-1]) < 5000


@pytest.mark.gpu
@pytest.mark.parametrize("sweeps", [1, 4, 7])
def test_gpu_fp64_sweeps_bit_exact(hip, oracle, sweeps):
    """Plain FP64 Jacobi sweeps on u (the solver's final polishing sweeps):
    bit-identical to the oracle's plain FP64 Jacobi, faces kept, for sizes
    from one cell to several tiles per axis."""
    rng = np.random.default_rng(60 + sweeps)
    sizes = [1, 5, 17, 33, 40]
    problems = [(n, rng.uniform(-1, 1, size=6), np.zeros((n + 2,) * 3, np.float32)) for n in sizes]
    for _, _, f in problems:
        f[1:-1, 1:-1, 1:-1] = rng.standard_normal(f[1:-1, 1:-1, 1:-1].shape)
    guesses = []
    for n, faces, _ in problems:
        g = padded(n, faces, oracle)
        g[1:-1, 1:-1, 1:-1] = rng.standard_normal((n, n, n))
        guesses.append(g)
    b = Batch(hip, problems, guesses)
    ub = hip.DeviceBuffer(b.elems * 8)
    host = np.zeros(b.elems)
    b.u.download(host)
    ub.upload(host)  # the same faces in both buffers
    in_b = C.c_int(-1)
    print("[case] poisson fp64 sweeps", sizes, sweeps, flush=True)
    hip.call("lgdhip_poisson_sweeps_f64", b.u.ptr, ub.ptr, b.f.ptr, b.probs, b.count, sweeps, b.ws.ptr, b.ws_bytes,
             C.byref(in_b), None)
    assert in_b.value == sweeps % 2
    out = np.empty(b.elems)
    (ub if in_b.value else b.u).download(out)
    for k, (n, _, f) in enumerate(problems):
        exp = guesses[k].copy()
        oracle.poisson_jacobi(exp, f, 1e-300, max_sweeps=sweeps, check=sweeps)
        o = b.probs[k].offset
        assert np.array_equal(out[o:o + exp.size].reshape(exp.shape), exp), n


@pytest.mark.gpu
def test_gpu_solve_with_fp64_polishing(hip, oracle):
    """final_fp64_sweeps = 8: the solve ends with 8 FP64 Jacobi sweeps on u --
    the polished u equals the unpolished solve's u plus 8 oracle FP64 sweeps,
    bit for bit; the reported residual is the polished one (recomputed on the
    CPU) and the sweep counts include the 8."""
    problems = synthetic.poisson_ensemble(5, sizes=(8, 16, 24), seed=12)
    plain, polished = Batch(hip, problems), Batch(hip, problems)
    r0 = plain.solve(rel_tol=1e-6, chunk=8)
    r1 = polished.solve(rel_tol=1e-6, chunk=8, polish=7)   # rounded up to 8
    u0, u1 = plain.u_blocks(), polished.u_blocks()
    for k, (n, faces, f) in enumerate(problems):
        exp = u0[k].copy()
        oracle.poisson_jacobi(exp, f, 1e-300, max_sweeps=8, check=8)
        assert np.array_equal(u1[k], exp), n
        assert r1[k].sweeps == r0[k].sweeps + 8
        check_solution(oracle, problems[k], u1[k], r1[k])
        assert r1[k].residual_norm <= r0[k].residual_norm
