"""Learned initial guess: the Poisson <-> U-Net bridge kernels and the
trained network's gain over the zero guess."""
import ctypes as C

import numpy as np
import pytest

from libgeodecomp_amd.unet import bf16_round

PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def sym_source(t, n):
    """Source coordinates (sx, sy, sz), each [n,n,n] indexed [z,y,x], of the
    output voxels under cube symmetry t: s[perm[i]] = o[i] with o = (x, y, z),
    then s[a] = n-1-s[a] for each set bit a of t // 6."""
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    o = (x, y, z)
    s = [None, None, None]
    for i, a in enumerate(PERMS[t % 6]):
        s[a] = o[i]
    for a in range(3):
        if (t // 6) >> a & 1:
            s[a] = n - 1 - s[a]
    return s


def face_field(n, faces):
    b = np.zeros((n, n, n))
    b[:, :, 0] += faces[0]
    b[:, :, -1] += faces[1]
    b[:, 0, :] += faces[2]
    b[:, -1, :] += faces[3]
    b[0, :, :] += faces[4]
    b[-1, :, :] += faces[5]
    return b


def apply_A0(u):
    """6u - sum_nb u with zero outside the cube."""
    p = np.pad(u, 1)
    return 6 * u - (p[:-2, 1:-1, 1:-1] + p[2:, 1:-1, 1:-1] + p[1:-1, :-2, 1:-1] + p[1:-1, 2:, 1:-1]
                    + p[1:-1, 1:-1, :-2] + p[1:-1, 1:-1, 2:])


@pytest.mark.parametrize("t", [0, 1, 2, 3, 4, 5, 6, 13, 29, 47])
def test_symmetry_maps_problems_to_problems(t):
    """u solves A0 u = f + b(faces) (FP64, manufactured); its image under
    symmetry t, with f and b read through the same symmetry, solves the
    same equation.  Pins the index map shared by input and target."""
    n = 5
    rng = np.random.default_rng(t)
    u = rng.standard_normal((n, n, n))
    faces = rng.uniform(-1, 1, 6)
    b = face_field(n, faces)
    f = apply_A0(u) - b
    sx, sy, sz = sym_source(t, n)
    u2, f2, b2 = u[sz, sy, sx], f[sz, sy, sx], b[sz, sy, sx]
    assert np.allclose(apply_A0(u2), f2 + b2, atol=1e-12)
    # the map is a bijection of the cube
    flat = (sz * n + sy) * n + sx
    assert np.array_equal(np.sort(flat.ravel()), np.arange(n ** 3))


def H_maps(H, n):
    """The six face basis fields [z,y,x] read from H (face x=-1 basis, H[z,y,x])
    at (x,y,z), (n-1-x,y,z), (y,x,z), (n-1-y,x,z), (z,y,x), (n-1-z,y,x)."""
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    at = lambda a, b, c: H[c, b, a]
    return [at(x, y, z), at(n - 1 - x, y, z), at(y, x, z), at(n - 1 - y, x, z), at(z, y, x), at(n - 1 - z, y, x)]


def face_solution(H, n, faces, sx=None, sy=None, sz=None):
    maps = H_maps(H.astype(np.float64), n)
    h = faces[0] * maps[0]
    for j in range(1, 6):
        h = h + faces[j] * maps[j]
    if sx is not None:
        h = h[sz, sy, sx]
    return h


def dense_laplace_solve(n, faces):
    """Exact solution of A0 u = b(faces) (f = 0) by a dense solve."""
    N = n ** 3
    A = np.zeros((N, N))
    idx = np.arange(N).reshape(n, n, n)
    for z in range(n):
        for y in range(n):
            for x in range(n):
                i = idx[z, y, x]
                A[i, i] = 6.0
                for dz, dy, dx in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    zz, yy, xx = z + dz, y + dy, x + dx
                    if 0 <= zz < n and 0 <= yy < n and 0 <= xx < n:
                        A[i, idx[zz, yy, xx]] = -1.0
    return np.linalg.solve(A, face_field(n, faces).ravel()).reshape(n, n, n)


def test_harmonic_basis_maps_give_the_face_solution():
    """sum_j face_j H_j with the six index maps equals the exact solution of
    the f = 0 problem with arbitrary faces (dense solve, n = 4)."""
    n = 4
    H = dense_laplace_solve(n, [1.0, 0, 0, 0, 0, 0])
    rng = np.random.default_rng(8)
    for _ in range(3):
        faces = rng.uniform(-1, 1, 6)
        assert np.allclose(face_solution(H, n, faces), dense_laplace_solve(n, faces), atol=1e-12)


@pytest.mark.gpu
def test_bridge_kernels_match_numpy(hip):
    from libgeodecomp_amd import synthetic
    from libgeodecomp_amd.learned_guess import Ensemble

    n = 8
    problems = synthetic.poisson_ensemble(2, sizes=(n,), seed=5)
    ens = Ensemble(problems)
    rng = np.random.default_rng(0)
    u_host = np.zeros(ens.elems)
    for k in range(2):
        o = ens.probs[k].offset
        u_host[o:o + (n + 2) ** 3] = rng.standard_normal((n + 2) ** 3)
    ens.u.upload(u_host)
    H = rng.uniform(0, 1, (n, n, n)).astype(np.float32)
    Hd = hip.DeviceBuffer(H.nbytes)
    Hd.upload(H)
    x = hip.DeviceBuffer(n ** 3 * 8 * 2)
    t = hip.DeviceBuffer(n ** 3 * 4)
    for k in range(2):
        fk = problems[k][2][1:-1, 1:-1, 1:-1]
        faces = np.asarray(problems[k][1], dtype=np.float64)
        bk = face_field(n, faces)
        uk = u_host[ens.probs[k].offset:ens.probs[k].offset + (n + 2) ** 3].reshape((n + 2,) * 3)[1:-1, 1:-1, 1:-1]
        for sym, inv, harm in ((0, 0.5, False), (7, -0.25, True), (47, 1.0 / 3.0, True), (20, 2.0, False)):
            hp = Hd.ptr if harm else None
            hip.call("lgdhip_poisson_unet_input", ens.f.ptr, C.byref(ens.probs[k]), inv, sym, hp, x.ptr, 8, None)
            hip.call("lgdhip_poisson_unet_target", ens.u.ptr, C.byref(ens.probs[k]), inv, sym, hp, t.ptr, None)
            xh = np.empty(n ** 3 * 8, np.uint16)
            x.download(xh)
            th = np.empty(n ** 3, np.float32)
            t.download(th)
            sx, sy, sz = sym_source(sym, n)
            xs = (xh.astype(np.uint32) << 16).view(np.float32).reshape(n, n, n, 8)
            assert np.array_equal(xs[..., 0], bf16_round((fk[sz, sy, sx].astype(np.float64) * inv).astype(np.float32)))
            assert np.array_equal(xs[..., 1], bf16_round((bk[sz, sy, sx] * inv).astype(np.float32)))
            h = face_solution(H, n, faces, sx, sy, sz) if harm else 0.0
            c = 2
            if harm:
                assert np.array_equal(xs[..., 2], bf16_round((h * inv).astype(np.float32)))
                c = 3
            assert np.all(xs[..., c:] == 0)
            assert np.array_equal(th.reshape(n, n, n), ((uk[sz, sy, sx] - h) * inv).astype(np.float32))
    # guesses: problem 0 with the harmonic part, problem 1 without; then the
    # face-only guess into a copy of problem 0's interior check
    pred = [rng.standard_normal(n ** 3).astype(np.float32) for _ in range(2)]
    for k in range(2):
        pb = hip.DeviceBuffer(pred[k].nbytes)
        pb.upload(pred[k])
        hip.call("lgdhip_poisson_unet_guess", pb.ptr, C.byref(ens.probs[k]), 3.0, Hd.ptr if k == 0 else None,
                 ens.u.ptr, None)
    uh = np.empty(ens.elems)
    ens.u.download(uh)
    for k in range(2):
        o = ens.probs[k].offset
        blk = uh[o:o + (n + 2) ** 3].reshape((n + 2,) * 3)
        ref = u_host[o:o + (n + 2) ** 3].reshape((n + 2,) * 3)
        inner = np.zeros_like(blk, bool)
        inner[1:-1, 1:-1, 1:-1] = True
        assert np.array_equal(blk[~inner], ref[~inner])      # halos untouched
        exp = pred[k].astype(np.float64).reshape(n, n, n) * 3.0
        if k == 0:
            exp = exp + face_solution(H, n, np.asarray(problems[0][1], dtype=np.float64))
        assert np.array_equal(blk[1:-1, 1:-1, 1:-1], exp)
    hip.call("lgdhip_poisson_face_guess", Hd.ptr, C.byref(ens.probs[1]), ens.u.ptr, None)
    ens.u.download(uh)
    o = ens.probs[1].offset
    blk = uh[o:o + (n + 2) ** 3].reshape((n + 2,) * 3)
    assert np.array_equal(blk[1:-1, 1:-1, 1:-1], face_solution(H, n, np.asarray(problems[1][1], dtype=np.float64)))


@pytest.mark.gpu
def test_smoothing_phase_damps_high_frequency_guess_error(hip):
    """A guess = solution + checkerboard noise: plain Jacobi damps the
    checkerboard as slowly as the smoothest mode, weighted Jacobi (2/3) by 3x
    per sweep, so the smoothed solve needs far fewer sweeps; both reach the
    tolerance relative to the zero guess's residual."""
    from libgeodecomp_amd import synthetic
    from libgeodecomp_amd.learned_guess import Ensemble

    n = 32
    ens = Ensemble(synthetic.poisson_ensemble(2, sizes=(n,), seed=9))
    ens.solve(1e-9)
    sols = [ens.interior(k) for k in range(2)]
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    cb = np.where((x + y + z) % 2 == 0, 1.0, -1.0)
    sweeps = {}
    for smooth in (0, 64):
        ens.reset()
        uh = np.empty(ens.elems)
        ens.u.download(uh)
        for k in range(2):
            o = ens.probs[k].offset
            blk = uh[o:o + (n + 2) ** 3].reshape((n + 2,) * 3)
            blk[1:-1, 1:-1, 1:-1] = sols[k] + 1e-3 * np.abs(sols[k]).max() * cb
        ens.u.upload(uh)
        res = ens.solve(1e-6, chunk=16, smooth_sweeps=smooth)
        assert all(r.converged == 1 for r in res)
        assert all(r.residual_norm <= 1e-6 * r.reference_norm for r in res)
        sweeps[smooth] = [r.sweeps for r in res]
    print("[smooth] sweeps without / with smoothing:", sweeps, flush=True)
    assert all(b < 0.5 * a for a, b in zip(sweeps[0], sweeps[64]))


@pytest.mark.gpu
def test_trained_network_beats_zero_guess(hip):
    """End to end at 32^3: solutions generated by the batched Jacobi solver,
    the U-Net trained on them by this library's kernels, then held-out
    problems solved from the zero guess and from the learned guess to the
    same 1e-6 residual (relative to the zero guess's residual).  The learned
    guess must be within 10 % of the solution and save >= 15 % of the sweeps
    (a ~3 % guess error saves ~30 %; sweeps scale with ln(1/tol) versus
    ln(err/tol) for the smooth error left after the smoothing phase)."""
    from libgeodecomp_amd import synthetic
    from libgeodecomp_amd.learned_guess import Ensemble, LearnedGuess, evaluate, harmonic_basis

    n = 32
    H = harmonic_basis(n)
    train = Ensemble(synthetic.poisson_ensemble(64, sizes=(n,), seed=101))
    res = train.solve(1e-6)
    assert all(r.converged for r in res)
    model = LearnedGuess(n, seed=7, harmonic=H)
    log = model.train(train, 600, lr=1e-3, batch=2, log_every=100)
    print("[learned] loss", log[0][1], "->", log[-1][1], flush=True)
    assert log[-1][1] < 0.1 * log[0][1]
    ev = Ensemble(synthetic.poisson_ensemble(8, sizes=(n,), seed=202))
    out = evaluate(model, ev, 1e-6, smooth_sweeps=64)
    arms = {a["arm"]: a for a in out["arms"]}
    for a in out["arms"]:
        print(f"[learned] {a['arm']}: sweeps {a['sweeps_mean']:.0f} guess err {a.get('guess_rel_error_median')}",
              flush=True)
        assert a["converged"] == 8 and a["final_rel_residual_max"] <= 1e-6
    assert arms["learned"]["guess_rel_error_median"] < 0.1
    assert arms["learned"]["sweeps_total"] <= 0.85 * arms["zero"]["sweeps_total"]
