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


# ----------------------------------------------------------------------
# coarse grid: numpy restatement of the restriction and prolongation
# ----------------------------------------------------------------------
def hat_weights(n, m):
    """W[j, i] = max(0, 1 - |i - X_j| / H), X_j = (j+1) H - 1, H = (n+1)/(m+1)."""
    H = (n + 1) / (m + 1)
    X = (np.arange(m) + 1) * H - 1.0
    return np.maximum(0.0, 1.0 - np.abs(np.arange(n)[None, :] - X[:, None]) / H), H


def restrict_np(f, m):
    """f [n,n,n] interior -> f_c [m,m,m] = H^2 x hat-weighted average."""
    n = f.shape[0]
    W, H = hat_weights(n, m)
    num = np.einsum("Zz,Yy,Xx,zyx->ZYX", W, W, W, f.astype(np.float64), optimize=True)
    den = np.einsum("Z,Y,X->ZYX", W.sum(1), W.sum(1), W.sum(1))
    return H * H * num / den


def pad_coarse(U, faces):
    """Interior [m,m,m] -> padded [m+2]^3 with faces, edges/corners = mean of the faces meeting there."""
    m = U.shape[0]
    P = np.zeros((m + 2,) * 3)
    P[1:-1, 1:-1, 1:-1] = U
    for z in range(m + 2):
        for y in range(m + 2):
            for x in range(m + 2):
                hs = []
                if x == 0: hs.append(faces[0])
                if x == m + 1: hs.append(faces[1])
                if y == 0: hs.append(faces[2])
                if y == m + 1: hs.append(faces[3])
                if z == 0: hs.append(faces[4])
                if z == m + 1: hs.append(faces[5])
                if hs:
                    P[z, y, x] = np.mean(hs)
    return P


def prolong_np(P, n):
    """Padded coarse [m+2]^3 -> fine [n,n,n], trilinear at padded coordinate (x+1)/H."""
    m = P.shape[0] - 2
    H = (n + 1) / (m + 1)
    I = np.zeros((n, m + 2))
    for x in range(n):
        xi = (x + 1) / H
        k = min(max(int(np.floor(xi)), 0), m)
        t = xi - k
        I[x, k] += 1 - t
        I[x, k + 1] += t
    return np.einsum("zc,yb,xa,cba->zyx", I, I, I, P, optimize=True)


def coarse_solve_np(fc, faces, H):
    """Dense solve of 6U - sum U_nb = H^2 fc (+ face terms) on the m^3 coarse grid."""
    m = fc.shape[0]
    N = m ** 3
    A = 6.0 * np.eye(N)
    idx = np.arange(N).reshape(m, m, m)
    for a in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[a], hi[a] = slice(0, m - 1), slice(1, m)
        A[idx[tuple(lo)].ravel(), idx[tuple(hi)].ravel()] = -1.0
        A[idx[tuple(hi)].ravel(), idx[tuple(lo)].ravel()] = -1.0
    rhs = fc + face_field(m, faces)
    return np.linalg.solve(A, rhs.ravel()).reshape(m, m, m)


def test_restriction_closed_forms():
    """Constant f -> H^2 c everywhere; f linear in x -> H^2 f(X_j) where the
    hat window lies inside the cube (with integer H the sampled hat weights
    are symmetric about X_j; otherwise only the average's smoothing holds)."""
    n, m = 31, 7
    W, H = hat_weights(n, m)
    assert np.allclose(restrict_np(np.full((n, n, n), 2.5), m), H * H * 2.5)
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    fc = restrict_np(3.0 + 0.5 * x, m)
    X = (np.arange(m) + 1) * H - 1.0
    inside = (X - H >= 0) & (X + H <= n - 1)
    assert inside.sum() >= 3
    assert np.allclose(fc[:, :, inside], H * H * (3.0 + 0.5 * X[inside])[None, None, :])


def test_prolongation_closed_forms():
    """Constant coarse field with equal faces -> that constant; a field
    linear along x (faces chosen on the line) -> the line at every fine cell
    whose stencil stays inside in y and z; coarse cell centres reproduced."""
    n, m = 40, 9
    H = (n + 1) / (m + 1)
    P = pad_coarse(np.full((m, m, m), 1.5), [1.5] * 6)
    assert np.allclose(prolong_np(P, n), 1.5)
    X = np.arange(-1, m + 1) * H + H - 1.0         # positions of padded indices 0..m+1
    line = 2.0 - 0.25 * X
    P = np.broadcast_to(line[None, None, :], (m + 2,) * 3).copy()
    out = prolong_np(P, n)
    xs = np.arange(n)
    yz = [i for i in range(n) if H <= i + 1 <= m * H]   # padded coordinate within [1, m]
    assert np.allclose(out[np.ix_(yz, yz, xs)], (2.0 - 0.25 * xs)[None, None, :])


def test_coarse_solution_approximates_the_fine_solution(oracle):
    """Restriction (with the H^2 scaling), a coarse solve with the fine faces
    and prolongation give the fine solution to the coarse grid's
    discretisation error: a wrong scale, spacing or boundary placement gives
    an O(1) error instead."""
    from libgeodecomp_amd import synthetic

    n, m = 32, 7
    (_, faces, fpad), = synthetic.poisson_ensemble(1, sizes=(n,), seed=4)
    u = np.zeros((n + 2,) * 3)
    oracle.poisson_faces(u, faces)
    sweeps, _ = oracle.poisson_jacobi(u, fpad, 1e-10)
    fine = u[1:-1, 1:-1, 1:-1]
    H = (n + 1) / (m + 1)
    Uc = coarse_solve_np(restrict_np(fpad[1:-1, 1:-1, 1:-1], m), faces, H)
    approx = prolong_np(pad_coarse(Uc, faces), n)
    err = np.linalg.norm(approx - fine) / np.linalg.norm(fine)
    print(f"[coarse] relative error of the prolongated coarse solution: {err:.3e}", flush=True)
    assert err < 0.15


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
    base = rng.uniform(-1, 1, (n, n, n)).astype(np.float32)
    bd = hip.DeviceBuffer(base.nbytes)
    bd.upload(base)
    x = hip.DeviceBuffer(n ** 3 * 8 * 2)
    t = hip.DeviceBuffer(n ** 3 * 4)
    for k in range(2):
        fk = problems[k][2][1:-1, 1:-1, 1:-1]
        faces = np.asarray(problems[k][1], dtype=np.float64)
        bk = face_field(n, faces)
        uk = u_host[ens.probs[k].offset:ens.probs[k].offset + (n + 2) ** 3].reshape((n + 2,) * 3)[1:-1, 1:-1, 1:-1]
        for sym, inv, use in ((0, 0.5, False), (7, -0.25, True), (47, 1.0 / 3.0, True), (20, 2.0, False)):
            bp = bd.ptr if use else None
            hip.call("lgdhip_poisson_unet_input", ens.f.ptr, C.byref(ens.probs[k]), inv, sym, bp, x.ptr, 8, None)
            hip.call("lgdhip_poisson_unet_target", ens.u.ptr, C.byref(ens.probs[k]), inv, sym, bp, t.ptr, None)
            xh = np.empty(n ** 3 * 8, np.uint16)
            x.download(xh)
            th = np.empty(n ** 3, np.float32)
            t.download(th)
            sx, sy, sz = sym_source(sym, n)
            xs = (xh.astype(np.uint32) << 16).view(np.float32).reshape(n, n, n, 8)
            assert np.array_equal(xs[..., 0], bf16_round((fk[sz, sy, sx].astype(np.float64) * inv).astype(np.float32)))
            assert np.array_equal(xs[..., 1], bf16_round((bk[sz, sy, sx] * inv).astype(np.float32)))
            bs = base[sz, sy, sx].astype(np.float64) if use else 0.0
            c = 2
            if use:
                assert np.array_equal(xs[..., 2], bf16_round((bs * inv).astype(np.float32)))
                c = 3
            assert np.all(xs[..., c:] == 0)
            assert np.array_equal(th.reshape(n, n, n), ((uk[sz, sy, sx] - bs) * inv).astype(np.float32))
    # guesses: problem 0 = pred * 3 + base, problem 1 = base alone
    pred = rng.standard_normal(n ** 3).astype(np.float32)
    pb = hip.DeviceBuffer(pred.nbytes)
    pb.upload(pred)
    hip.call("lgdhip_poisson_unet_guess", pb.ptr, C.byref(ens.probs[0]), 3.0, bd.ptr, ens.u.ptr, None)
    hip.call("lgdhip_poisson_unet_guess", None, C.byref(ens.probs[1]), 1.0, bd.ptr, ens.u.ptr, None)
    uh = np.empty(ens.elems)
    ens.u.download(uh)
    expect = [pred.astype(np.float64).reshape(n, n, n) * 3.0 + base.astype(np.float64), base.astype(np.float64)]
    for k in range(2):
        o = ens.probs[k].offset
        blk = uh[o:o + (n + 2) ** 3].reshape((n + 2,) * 3)
        ref = u_host[o:o + (n + 2) ** 3].reshape((n + 2,) * 3)
        inner = np.zeros_like(blk, bool)
        inner[1:-1, 1:-1, 1:-1] = True
        assert np.array_equal(blk[~inner], ref[~inner])      # halos untouched
        assert np.array_equal(blk[1:-1, 1:-1, 1:-1], expect[k])


@pytest.mark.gpu
@pytest.mark.parametrize("n,m", [(32, 7), (40, 9), (17, 4), (128, 31)])
def test_coarse_restrict_prolong_match_numpy(hip, n, m):
    from libgeodecomp_amd import synthetic
    from libgeodecomp_amd.learned_guess import Ensemble

    problems = synthetic.poisson_ensemble(2, sizes=(n,), seed=n + m)
    fine = Ensemble(problems)
    zeros = np.zeros((m + 2,) * 3, np.float32)
    coarse = Ensemble([(m, p[1], zeros) for p in problems])
    assert hip.load().lgdhip_poisson_coarse_spacing(n, m) == (n + 1) / (m + 1)
    rng = np.random.default_rng(m)
    uc = np.zeros(coarse.elems)
    for k in range(2):
        hip.call("lgdhip_poisson_coarse_restrict", fine.f.ptr, C.byref(fine.probs[k]), coarse.f.ptr,
                 C.byref(coarse.probs[k]), None)
        o = coarse.probs[k].offset
        P = pad_coarse(rng.standard_normal((m, m, m)), problems[k][1])
        uc[o:o + (m + 2) ** 3] = P.ravel()
    # the device reads face cells from the pool and edge/corner cells from the faces
    coarse.u.upload(uc)
    fch = np.empty(coarse.elems, np.float32)
    coarse.f.download(fch)
    out = hip.DeviceBuffer(n ** 3 * 4)
    for k in range(2):
        o = coarse.probs[k].offset
        got = fch[o:o + (m + 2) ** 3].reshape((m + 2,) * 3)
        exp = restrict_np(problems[k][2][1:-1, 1:-1, 1:-1], m)
        scale = np.abs(exp).max()
        assert np.allclose(got[1:-1, 1:-1, 1:-1], exp, rtol=1e-6, atol=1e-6 * scale)
        hip.call("lgdhip_poisson_coarse_prolong", coarse.u.ptr, C.byref(coarse.probs[k]), n, out.ptr, None)
        oh = np.empty(n ** 3, np.float32)
        out.download(oh)
        P = uc[o:o + (m + 2) ** 3].reshape((m + 2,) * 3)
        expp = prolong_np(P, n)
        assert np.allclose(oh.reshape(n, n, n), expp, rtol=1e-6, atol=1e-6 * np.abs(expp).max())


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
def test_trained_network_beats_zero_guess(hip, tmp_path):
    """End to end at 32^3: solutions generated by the batched Jacobi solver,
    the U-Net trained on them by this library's kernels on top of the
    prolongated 7^3 coarse solution, then held-out problems solved from the
    zero guess, from the coarse solution alone and from the learned guess to
    the same 1e-6 residual (relative to the zero guess's residual).  The
    learned guess must be closer to the solution than the coarse one and
    save >= 25 % of the zero guess's sweeps."""
    from libgeodecomp_amd import synthetic
    from libgeodecomp_amd.learned_guess import CoarseBase, Ensemble, LearnedGuess, evaluate

    n, m = 32, 7
    train = Ensemble(synthetic.poisson_ensemble(64, sizes=(n,), seed=101))
    res = train.solve(1e-6)
    assert all(r.converged for r in res)
    base = CoarseBase(train, m)
    base.compute()
    model = LearnedGuess(n, seed=7)
    log = model.train(train, base, 600, lr=1e-3, batch=2, log_every=100)
    print("[learned] loss", log[0][1], "->", log[-1][1], flush=True)
    assert log[-1][1] < 0.5 * log[0][1]
    ev = Ensemble(synthetic.poisson_ensemble(8, sizes=(n,), seed=202))
    out = evaluate(model, ev, m, 1e-6, smooth_sweeps=64)
    arms = {a["arm"]: a for a in out["arms"]}
    for a in out["arms"]:
        print(f"[learned] {a['arm']}: sweeps {a['sweeps_mean']:.0f} guess err {a.get('guess_rel_error_median')}",
              flush=True)
        assert a["converged"] == 8 and a["final_rel_residual_max"] <= 1e-6
    assert arms["learned"]["guess_rel_error_median"] < arms["coarse"]["guess_rel_error_median"]
    assert arms["learned"]["sweeps_total"] <= 0.75 * arms["zero"]["sweeps_total"]
    # saved and loaded (the bench's form of a trained network): the same BF16
    # weights, so the same guess bit for bit
    path = tmp_path / "net.npz"
    model.save(path, {"train_steps": 600})
    loaded, meta = LearnedGuess.load(path)
    assert meta == {"train_steps": 600}
    bev = CoarseBase(ev, m)
    bev.compute()
    s = LearnedGuess.scales(ev)
    guesses = []
    for mdl in (model, loaded):
        ev.reset()
        mdl.propose(ev, bev, 3, s[3])
        u = np.empty(ev.elems, np.float64)
        ev.u.download(u)
        guesses.append(u)
    assert np.array_equal(guesses[0], guesses[1])
