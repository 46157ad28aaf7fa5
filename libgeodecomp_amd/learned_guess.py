"""Learned initial guess for the Poisson ensemble.

Plain Jacobi spends its sweeps on the smoothest error modes (the iteration
damps them by cos(pi/(n+1)) per sweep), and those are the global modes a
4-level U-Net's receptive field does not span at 128^3.  So the guess is
built in two parts:

  * base: the problem restricted to a coarse grid (m^3, spacing
    H = (n+1)/(m+1), same faces), solved there by the same batched Jacobi
    solver (cheap: m^3 cells, ~m^2 sweeps), prolongated trilinearly -- it
    carries the smooth modes to the coarse grid's discretisation error;
  * learned correction: the 3D U-Net (libgeodecomp_amd.unet) maps
    (f, b, base) / s to (u - base) / s, trained on steady states that the
    batched Jacobi solver generates, every problem read through one of the 48
    symmetries of the cube and a random sign (exact augmentations), with
    bias-free layers (a positively homogeneous network) and per-sample
    normalised targets (every sample weighs by its relative error).

The guess u0 = base + s * net(...) is then polished by the same solver -- a
short weighted-Jacobi smoothing phase for the high-frequency part of the
network's error, then plain Jacobi -- to the same relative residual as the
plain solve from the zero guess (both relative to the zero guess's residual
||f + b||_2).  evaluate() reports the zero guess, the base alone and the
learned guess, so the network's own contribution is visible.

No reference counterpart.  Everything on the device is a liblgdhip kernel;
this module only orchestrates (and times) them.

    python -m libgeodecomp_amd.learned_guess --n 128 --train 512 --eval 16 --steps 30000 --lr 5e-4
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import sys
import time

import numpy as np

from . import lgdhip, synthetic
from .unet import UNet3D, from_bf16_bits


def _sync():
    lgdhip.call("lgdhip_device_synchronize")


def default_coarse_size(n: int, spacing: float = 4.0) -> int:
    """Coarse cells per edge for a coarse spacing of about `spacing` fine cells."""
    return max(1, int(round((n + 1) / spacing)) - 1)


class Ensemble:
    """n^3 Poisson problems (one size) in device pools: f (FP32), u (FP64)."""

    def __init__(self, problems, stream=None):
        self.lib = lgdhip.load()
        self.count = len(problems)
        self.n = problems[0][0]
        assert all(p[0] == self.n for p in problems), "one cube size per ensemble"
        self.stream = stream
        self.probs = (lgdhip.lgdhip_poisson_problem * self.count)()
        self.fmax = np.zeros(self.count)
        self.facemax = np.zeros(self.count)
        for k, (n, faces, f) in enumerate(problems):
            self.probs[k].n = n
            for j in range(6):
                self.probs[k].face[j] = faces[j]
            self.fmax[k] = float(np.abs(f).max())
            self.facemax[k] = float(np.abs(np.asarray(faces)).max())
        self.elems = self.lib.lgdhip_poisson_layout(self.probs, self.count)
        self.ws_bytes = self.lib.lgdhip_poisson_workspace_bytes(self.probs, self.count)
        self.ws = lgdhip.DeviceBuffer(self.ws_bytes)
        self.u = lgdhip.DeviceBuffer(self.elems * 8)
        self.f = lgdhip.DeviceBuffer(self.elems * 4)
        block = (self.n + 2) ** 3
        for k, (_, _, f) in enumerate(problems):
            host = np.ascontiguousarray(f, dtype=np.float32).ravel()
            assert host.size == block
            dst = C.c_void_p(self.f.value + self.probs[k].offset * 4)
            lgdhip.call("lgdhip_memcpy_h2d", dst, host.ctypes.data_as(C.c_void_p), host.nbytes, self.stream)
        self.reset()
        self.reference = self.residual_norms()

    def prob(self, k: int):
        return C.byref(self.probs[k])

    def reset(self):
        """u = 0 in every interior, faces in every halo."""
        lgdhip.call("lgdhip_fill_u8", self.u.ptr, 0, self.elems * 8, self.stream)
        lgdhip.call("lgdhip_poisson_set_faces", self.u.ptr, self.probs, self.count, self.ws.ptr, self.ws_bytes,
                    self.stream)

    def residual_norms(self) -> np.ndarray:
        out = np.zeros(self.count)
        lgdhip.call("lgdhip_poisson_residual_norms", self.u.ptr, self.f.ptr, self.probs, self.count, self.ws.ptr,
                    self.ws_bytes, out.ctypes.data_as(C.POINTER(C.c_double)), self.stream)
        return out

    def solve(self, rel_tol: float = 1e-6, chunk: int = 64, use_graph: int = 1, max_sweeps: int = 10_000_000,
              smooth_sweeps: int = 0, smooth_omega: float = 2.0 / 3.0, polish: int = 0):
        """Solves every problem from the current u to rel_tol relative to the
        zero guess's residual norm, then `polish` plain FP64 Jacobi sweeps on
        u; returns the per-problem results."""
        opt = lgdhip.lgdhip_poisson_options()
        opt.rel_tol, opt.chunk, opt.max_sweeps, opt.use_graph = rel_tol, chunk, max_sweeps, use_graph
        opt.smooth_sweeps, opt.smooth_omega = smooth_sweeps, smooth_omega
        opt.final_fp64_sweeps = polish
        ref = (C.c_double * self.count)(*self.reference)
        opt.reference_norms = ref
        res = (lgdhip.lgdhip_poisson_result * self.count)()
        lgdhip.call("lgdhip_poisson_batch_solve", self.u.ptr, self.f.ptr, self.probs, self.count, C.byref(opt),
                    self.ws.ptr, self.ws_bytes, res, self.stream)
        return res

    def interior(self, k: int) -> np.ndarray:
        """Problem k's interior of u (FP64, [z, y, x])."""
        n = self.n
        host = np.empty((n + 2) ** 3)
        src = C.c_void_p(self.u.value + self.probs[k].offset * 8)
        lgdhip.call("lgdhip_memcpy_d2h", host.ctypes.data_as(C.c_void_p), src, host.nbytes, self.stream)
        return host.reshape((n + 2,) * 3)[1:-1, 1:-1, 1:-1].copy()


class CoarseBase:
    """The base field of every problem of an ensemble: its coarse-grid
    solution (m^3, same faces) prolongated to the fine cells (dense n^3 FP32
    per problem)."""

    def __init__(self, data: Ensemble, m: int, rel_tol: float = 1e-6):
        self.data, self.m, self.rel_tol = data, m, rel_tol
        n = data.n
        zeros = np.zeros((m + 2,) * 3, np.float32)
        faces = [np.array([data.probs[k].face[j] for j in range(6)]) for k in range(data.count)]
        self.coarse = Ensemble([(m, faces[k], zeros) for k in range(data.count)], stream=data.stream)
        self.field = lgdhip.DeviceBuffer(data.count * n ** 3 * 4)
        self.results = None

    def ptr(self, k: int) -> C.c_void_p:
        return C.c_void_p(self.field.value + k * self.data.n ** 3 * 4)

    def compute(self):
        """Restrict, solve on the coarse grid, prolongate (all problems)."""
        d, c = self.data, self.coarse
        for k in range(d.count):
            lgdhip.call("lgdhip_poisson_coarse_restrict", d.f.ptr, d.prob(k), c.f.ptr, c.prob(k), d.stream)
        c.reset()
        c.reference = c.residual_norms()
        self.results = c.solve(self.rel_tol, chunk=32)
        for k in range(d.count):
            lgdhip.call("lgdhip_poisson_coarse_prolong", c.u.ptr, c.prob(k), d.n, self.ptr(k), d.stream)


class LearnedGuess:
    """U-Net proposing initial guesses for n^3 problems: (u - base)/s from
    (f, b, base)/s, s = max(|f|, |faces|); without a base, u/s from (f, b)/s."""

    def __init__(self, n: int, channels=(16, 32, 64, 128), seed: int = 0, use_base: bool = True, stream=None):
        self.n = n
        self.use_base = use_base
        self.stream = stream
        self.net = UNet3D(n=n, channels=channels, in_ch=3 if use_base else 2, seed=seed, stream=stream, bias=False)
        self.target = lgdhip.DeviceBuffer(n ** 3 * 4)
        self.rng = np.random.default_rng(seed + 1)

    @staticmethod
    def scales(data: Ensemble) -> np.ndarray:
        return np.maximum(np.maximum(data.fmax, data.facemax), 1e-30)

    def _base(self, base: CoarseBase | None, k: int):
        return base.ptr(k) if (self.use_base and base is not None) else None

    def _input(self, data: Ensemble, base, k: int, sym: int, inv: float):
        lgdhip.call("lgdhip_poisson_unet_input", data.f.ptr, data.prob(k), inv, sym, self._base(base, k),
                    self.net.x.ptr, self.net.x.ld, self.stream)

    def _target(self, data: Ensemble, base, k: int, sym: int, inv: float):
        lgdhip.call("lgdhip_poisson_unet_target", data.u.ptr, data.prob(k), inv, sym, self._base(base, k),
                    self.target.ptr, self.stream)

    def sample_norms(self, data: Ensemble, base) -> np.ndarray:
        """RMS of every problem's target at unit input scale (data.u must
        hold the solutions)."""
        s = self.scales(data)
        out = np.zeros(data.count)
        for k in range(data.count):
            self._target(data, base, k, 0, 1.0 / s[k])
            t = np.empty(self.n ** 3, np.float32)
            self.target.download(t, self.stream)
            out[k] = float(np.sqrt(np.mean(t.astype(np.float64) ** 2)))
        return out

    def train(self, data: Ensemble, base: CoarseBase | None, steps: int, lr: float = 1e-3, batch: int = 2,
              log_every: int = 0, lr_final: float = 0.05, verbose: bool = False) -> list:
        """Adam on mini-batches of `batch` augmented samples (gradients
        accumulated), cosine learning-rate decay to lr * lr_final.  data.u
        must hold the solutions.  Every sample is scaled to a unit-RMS
        target (exact for the bias-free network).  Returns (step, loss)."""
        assert data.n == self.n
        s = self.scales(data)
        rms = self.sample_norms(data, base)
        g = np.where(rms > 0, 1.0 / np.maximum(rms, 1e-300), 0.0)
        log = []
        for step in range(steps):
            for b in range(batch):
                k = int(self.rng.integers(data.count))
                sym = int(self.rng.integers(48))
                sign = 1.0 if self.rng.integers(2) else -1.0
                inv = sign * g[k] / s[k]
                self._input(data, base, k, sym, inv)
                self._target(data, base, k, sym, inv)
                self.net.forward()
                read = log_every and (step % log_every == 0 or step == steps - 1) and b == 0
                loss = self.net.loss(self.target, read=bool(read))
                if read:
                    log.append((step, loss))
                    if verbose:
                        print(f"[learned] step {step} loss {loss:.4e}", flush=True)
                self.net.backward(accumulate=1 if b else 0)
            frac = step / max(1, steps - 1)
            lr_t = lr * (lr_final + (1.0 - lr_final) * 0.5 * (1.0 + math.cos(math.pi * frac)))
            self.net.adam(lr=lr_t)
        return log

    def save(self, path, meta: dict | None = None):
        """The network's BF16 weights (the copies the forward pass reads) and
        its shape in an .npz file; `meta` (JSON-serialisable) is kept beside
        them (how the weights were trained)."""
        arrays = {}
        for name in self.net.order:
            l = self.net.layers[name]
            wb = np.empty(l.cout * l.kp, np.uint16)
            l.wb.download(wb, self.stream)
            arrays["w_" + name] = wb.reshape(l.cout, l.kp)
        info = {"n": self.n, "channels": list(self.net.ch), "use_base": self.use_base, "meta": meta or {}}
        np.savez_compressed(path, info=np.frombuffer(json.dumps(info).encode(), np.uint8), **arrays)

    @classmethod
    def load(cls, path, stream=None):
        """A LearnedGuess with the weights `save` wrote (bit-identical BF16
        weights, so the same guesses); returns (model, meta)."""
        with np.load(path, allow_pickle=False) as z:
            info = json.loads(bytes(z["info"]).decode())
            model = cls(info["n"], channels=tuple(info["channels"]), use_base=info["use_base"], stream=stream)
            for name in model.net.order:
                l = model.net.layers[name]
                taps = 27 if l.ksize == 3 else 1
                wb = z["w_" + name]
                if wb.shape != (l.cout, l.kp):
                    raise ValueError(f"{path}: layer {name} has shape {wb.shape}, expected {(l.cout, l.kp)}")
                w = from_bf16_bits(wb[:, :taps * l.cin]).reshape(l.cout, taps, l.cin)
                model.net.set_weights(name, w, np.zeros(l.cout, np.float32))
        return model, info["meta"]

    def propose(self, data: Ensemble, base: CoarseBase | None, k: int, s: float):
        """Writes the guess for problem k into data.u."""
        self._input(data, base, k, 0, 1.0 / s)
        pred = self.net.forward()
        lgdhip.call("lgdhip_poisson_unet_guess", pred.ptr, data.prob(k), s, self._base(base, k), data.u.ptr,
                    self.stream)


def _arm(name: str, exact: list, results, t_guess: float, t_solve: float, r_guess=None, guess=None) -> dict:
    sw = np.array([r.sweeps for r in results], dtype=np.float64)
    out = {"arm": name, "converged": int(sum(r.converged for r in results)), "sweeps_mean": float(sw.mean()),
           "sweeps_total": float(sw.sum()), "wall_guess_s": t_guess, "wall_solve_s": t_solve,
           "wall_s": t_guess + t_solve,
           "final_rel_residual_max": float(max(r.residual_norm / r.reference_norm for r in results))}
    if guess is not None:
        err = [float(np.linalg.norm(g - e) / max(np.linalg.norm(e), 1e-300)) for g, e in zip(guess, exact)]
        out["guess_rel_error_median"] = float(np.median(err))
        out["guess_rel_error_max"] = float(np.max(err))
        out["guess_rel_residual_median"] = float(np.median(r_guess))
    return out


def evaluate(model: LearnedGuess, data: Ensemble, coarse_m: int = 0, rel_tol: float = 1e-6, chunk: int = 64,
             smooth_sweeps: int = 128, polish: int = 16) -> dict:
    """Zero guess (plain Jacobi); with a coarse grid, the prolongated coarse
    solution alone; and the learned guess -- on the same problems with the
    same solver, tolerance and reference norms (the zero guess's residual),
    every arm ending with `polish` plain FP64 Jacobi sweeps on u.
    Guessed arms start with `smooth_sweeps` weighted (2/3) Jacobi sweeps,
    counted in their sweeps; their wall times include building the guess
    (restriction, coarse solve, prolongation, network)."""
    data.reset()
    _sync()
    t0 = time.perf_counter()
    res0 = data.solve(rel_tol, chunk, polish=polish)
    _sync()
    t_zero = time.perf_counter() - t0
    exact = [data.interior(k) for k in range(data.count)]
    arms = [_arm("zero", exact, res0, 0.0, t_zero)]
    base = CoarseBase(data, coarse_m) if coarse_m > 0 else None
    s = model.scales(data)

    def guessed(name, make):
        data.reset()
        _sync()
        t0 = time.perf_counter()
        make()
        _sync()
        t_g = time.perf_counter() - t0
        guess = [data.interior(k) for k in range(data.count)]
        r_guess = data.residual_norms() / data.reference
        t0 = time.perf_counter()
        res = data.solve(rel_tol, chunk, smooth_sweeps=smooth_sweeps, polish=polish)
        _sync()
        arms.append(_arm(name, exact, res, t_g, time.perf_counter() - t0, r_guess, guess))

    if base is not None:
        def coarse_only():
            base.compute()
            for k in range(data.count):
                lgdhip.call("lgdhip_poisson_unet_guess", None, data.prob(k), 1.0, base.ptr(k), data.u.ptr,
                            data.stream)
        guessed("coarse", coarse_only)

    def learned():
        if base is not None:
            base.compute()
        for k in range(data.count):
            model.propose(data, base, k, s[k])
    guessed("learned", learned)
    best = arms[-1]
    return {"problems": data.count, "n": data.n, "coarse_m": coarse_m, "rel_tol": rel_tol,
            "smooth_sweeps": smooth_sweeps, "fp64_polish_sweeps": polish, "arms": arms,
            "sweep_ratio": arms[0]["sweeps_total"] / max(best["sweeps_total"], 1.0),
            "wall_speedup": arms[0]["wall_s"] / best["wall_s"]}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--train", type=int, default=512, help="solved training problems")
    ap.add_argument("--eval", type=int, default=16, help="held-out problems")
    ap.add_argument("--steps", type=int, default=30000)
    ap.add_argument("--batch", type=int, default=2)
    ap.add_argument("--lr", type=float, default=5e-4, help="peak Adam learning rate (cosine decay)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--coarse", type=int, default=-1, help="coarse cells per edge (-1: spacing ~4, 0: no base)")
    ap.add_argument("--smooth", type=int, default=128, help="weighted-Jacobi sweeps before polishing a guess")
    ap.add_argument("--seed", type=int, default=synthetic.DEFAULT_SEED)
    ap.add_argument("--log-every", type=int, default=250)
    ap.add_argument("--save", default="", help="write the trained network's weights to this .npz file")
    ap.add_argument("--load", default="", help="evaluate the network in this .npz file (no generation, no training)")
    a = ap.parse_args(argv)
    lgdhip.load()
    lgdhip.call("lgdhip_set_device", 0)
    stream = lgdhip.Stream()
    m = default_coarse_size(a.n) if a.coarse < 0 else a.coarse
    out = {"n": a.n, "coarse_m": m, "train_problems": a.train, "eval_problems": a.eval, "steps": a.steps,
           "batch": a.batch}

    if a.load:
        model, meta = LearnedGuess.load(a.load, stream=stream.h)
        out["weights"] = {"file": a.load, **meta}
        ev = Ensemble(synthetic.poisson_ensemble(a.eval, sizes=(model.n,), seed=a.seed + 7919), stream=stream.h)
        out["eval"] = evaluate(model, ev, m, a.tol, smooth_sweeps=a.smooth)
        for arm in out["eval"]["arms"]:
            print(f"[learned] {arm['arm']:8s} sweeps {arm['sweeps_mean']:9.0f}  wall {arm['wall_s']:.3f} s", flush=True)
        print(json.dumps(out), flush=True)
        return out

    t0 = time.perf_counter()
    train = Ensemble(synthetic.poisson_ensemble(a.train, sizes=(a.n,), seed=a.seed), stream=stream.h)
    res = train.solve(a.tol)
    base = None
    if m > 0:
        base = CoarseBase(train, m)
        base.compute()
    _sync()
    out["generate_s"] = time.perf_counter() - t0
    out["generate_sweeps_mean"] = float(np.mean([r.sweeps for r in res]))
    out["generate_converged"] = int(sum(r.converged for r in res))
    print(f"[learned] generated {a.train} solutions of {a.n}^3 (+ {m}^3 coarse) in {out['generate_s']:.1f} s "
          f"({out['generate_sweeps_mean']:.0f} sweeps each)", flush=True)

    model = LearnedGuess(a.n, seed=a.seed % 1000, use_base=m > 0, stream=stream.h)
    t0 = time.perf_counter()
    log = model.train(train, base, a.steps, lr=a.lr, batch=a.batch, log_every=a.log_every, verbose=True)
    _sync()
    out["train_s"] = time.perf_counter() - t0
    out["loss_log"] = log
    print(f"[learned] trained {a.steps} steps in {out['train_s']:.1f} s, loss {log[0][1]:.3e} -> {log[-1][1]:.3e}",
          flush=True)
    del base, train
    if a.save:
        model.save(a.save, {"train_problems": a.train, "train_steps": a.steps, "batch": a.batch, "peak_lr": a.lr,
                            "coarse_m": m, "seed": a.seed, "final_loss": log[-1][1] if log else None,
                            "generate_s": round(out["generate_s"], 1), "train_s": round(out["train_s"], 1),
                            "command": "python -m libgeodecomp_amd.learned_guess " + " ".join(argv or sys.argv[1:])})
        print(f"[learned] weights written to {a.save}", flush=True)

    ev = Ensemble(synthetic.poisson_ensemble(a.eval, sizes=(a.n,), seed=a.seed + 7919), stream=stream.h)
    out["eval"] = evaluate(model, ev, m, a.tol, smooth_sweeps=a.smooth)
    for arm in out["eval"]["arms"]:
        print(f"[learned] {arm['arm']:8s} sweeps {arm['sweeps_mean']:9.0f}  wall {arm['wall_s']:.3f} s  "
              f"guess err {arm.get('guess_rel_error_median', 1.0):.2e}", flush=True)
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
