"""3D U-Net for the learned initial guess of the Poisson ensemble.

Host orchestration only: every tensor operation is a liblgdhip kernel
(MFMA BF16 convolutions with FP32 accumulation, pooling, upsampling, loss,
Adam), called through the C-ABI.  Architecture: 4 levels with 16-32-64-128
channels, two 3x3x3 convolutions + ReLU per level, 2x2x2 average pooling,
nearest upsampling with skip concatenation, a final 1x1x1 convolution to one
FP32 channel.  Inputs: the source f and the face field b (each cell's sum
of the Dirichlet values of the faces it touches -- exactly the boundary part
of the zero-guess residual), both divided by the sample scale s; output
u / (s * target_scale).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import lgdhip


# ----------------------------------------------------------------------
# BF16 helpers (round to nearest even, as the kernels)
# ----------------------------------------------------------------------
def to_bf16_bits(a: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(b: np.ndarray) -> np.ndarray:
    return (np.ascontiguousarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16_round(a: np.ndarray) -> np.ndarray:
    return from_bf16_bits(to_bf16_bits(a))


def face_field(n: int, faces) -> np.ndarray:
    """b[z,y,x] = sum of the Dirichlet values of the faces cell (x,y,z) touches."""
    b = np.zeros((n, n, n), dtype=np.float64)
    b[:, :, 0] += faces[0]
    b[:, :, -1] += faces[1]
    b[:, 0, :] += faces[2]
    b[:, -1, :] += faces[3]
    b[0, :, :] += faces[4]
    b[-1, :, :] += faces[5]
    return b


def conv_weights_layout(w: np.ndarray, kp: int) -> np.ndarray:
    """[cout, taps, cin] -> [cout, Kp] with k = tap*cin + ci, zero padded."""
    cout, taps, cin = w.shape
    out = np.zeros((cout, kp), dtype=w.dtype)
    out[:, :taps * cin] = w.reshape(cout, taps * cin)
    return out


# ----------------------------------------------------------------------
# device tensors
# ----------------------------------------------------------------------
class Tensor:
    """Channels-last BF16 cube n^3 x ld (owning)."""

    def __init__(self, n: int, ld: int):
        self.n, self.ld = n, ld
        self.nvox = n ** 3
        self.buf = lgdhip.DeviceBuffer(self.nvox * ld * 2)

    @property
    def ptr(self):
        return self.buf.ptr

    def zero(self, stream=None):
        lgdhip.call("lgdhip_fill_u8", self.ptr, 0, self.nvox * self.ld * 2, stream)

    def numpy(self) -> np.ndarray:
        h = np.empty(self.nvox * self.ld, dtype=np.uint16)
        self.buf.download(h)
        return from_bf16_bits(h).reshape(self.n, self.n, self.n, self.ld)


@dataclass
class Conv:
    name: str
    cin: int
    cout: int
    ksize: int
    relu: bool
    n: int
    kp: int = 0
    w32: object = None      # FP32 master [cout][kp]
    wb: object = None       # BF16 copy
    wt: object = None       # BF16 dgrad weights [cin][kp(cout)]
    bias: object = None     # FP32 [cout]
    gw: object = None       # FP32 gradient
    gb: object = None
    m: list = field(default_factory=list)
    v: list = field(default_factory=list)


class UNet3D:
    """Weights, activations and gradients of one U-Net on n^3 volumes."""

    def __init__(self, n: int = 128, channels=(16, 32, 64, 128), in_ch: int = 2, seed: int = 0, stream=None,
                 bias: bool = True):
        """bias=False drops every bias: the network is then positively
        homogeneous, net(c x) = c net(x) for c > 0."""
        assert n % 8 == 0 and n >= 8, "three 2x poolings need n to be a multiple of 8"
        assert 1 <= in_ch <= 8
        self.n, self.ch, self.in_ch = n, channels, in_ch
        self.stream = stream
        self.use_bias = bias
        self.lib = lgdhip.load()
        c0, c1, c2, c3 = channels
        s = [n, n // 2, n // 4, n // 8]
        L = []
        L.append(Conv("e0a", in_ch, c0, 3, True, s[0]))
        L.append(Conv("e0b", c0, c0, 3, True, s[0]))
        L.append(Conv("e1a", c0, c1, 3, True, s[1]))
        L.append(Conv("e1b", c1, c1, 3, True, s[1]))
        L.append(Conv("e2a", c1, c2, 3, True, s[2]))
        L.append(Conv("e2b", c2, c2, 3, True, s[2]))
        L.append(Conv("e3a", c2, c3, 3, True, s[3]))
        L.append(Conv("e3b", c3, c3, 3, True, s[3]))
        L.append(Conv("d2a", c3 + c2, c2, 3, True, s[2]))
        L.append(Conv("d2b", c2, c2, 3, True, s[2]))
        L.append(Conv("d1a", c2 + c1, c1, 3, True, s[1]))
        L.append(Conv("d1b", c1, c1, 3, True, s[1]))
        L.append(Conv("d0a", c1 + c0, c0, 3, True, s[0]))
        L.append(Conv("d0b", c0, c0, 3, True, s[0]))
        L.append(Conv("out", c0, 1, 1, False, s[0]))
        self.layers = {l.name: l for l in L}
        self.order = [l.name for l in L]
        rng = np.random.default_rng(seed)
        for l in L:
            l.kp = self.lib.lgdhip_conv3d_kpad(l.cin, l.ksize)
            taps = 27 if l.ksize == 3 else 1
            std = np.sqrt(2.0 / (taps * l.cin)) if l.relu else np.sqrt(1.0 / (taps * l.cin))
            w = rng.standard_normal((l.cout, taps, l.cin)) * std
            self._set_weights(l, w, np.zeros(l.cout))
            l.gw = lgdhip.DeviceBuffer(l.cout * l.kp * 4)
            l.gb = lgdhip.DeviceBuffer(l.cout * 4)
            l.m = [lgdhip.DeviceBuffer(l.cout * l.kp * 4), lgdhip.DeviceBuffer(l.cout * 4)]
            l.v = [lgdhip.DeviceBuffer(l.cout * l.kp * 4), lgdhip.DeviceBuffer(l.cout * 4)]
            for b in l.m + l.v + [l.gb]:
                lgdhip.call("lgdhip_fill_u8", b.ptr, 0, b.nbytes, stream)
        self.t = 0
        # activations (forward) and their gradients
        self.x = Tensor(s[0], 8)                     # input: channels 0..in_ch-1, padded to 8
        self.a0, self.cat0 = Tensor(s[0], c0), Tensor(s[0], c1 + c0)   # cat0 = [up(d1b) | e0b (= s0)]
        self.p1, self.a1, self.cat1 = Tensor(s[1], c0), Tensor(s[1], c1), Tensor(s[1], c2 + c1)
        self.p2, self.a2, self.cat2 = Tensor(s[2], c1), Tensor(s[2], c2), Tensor(s[2], c3 + c2)
        self.p3, self.a3, self.b3 = Tensor(s[3], c2), Tensor(s[3], c3), Tensor(s[3], c3)
        # the encoder's skip outputs as dense tensors: pooled from, and copied
        # into the concatenations by the upsampling as whole rows
        self.s0, self.s1, self.s2 = Tensor(s[0], c0), Tensor(s[1], c1), Tensor(s[2], c2)
        self.a2d, self.b2d = Tensor(s[2], c2), Tensor(s[2], c2)
        self.a1d, self.b1d = Tensor(s[1], c1), Tensor(s[1], c1)
        self.a0d, self.b0d = Tensor(s[0], c0), Tensor(s[0], c0)
        self.pred = lgdhip.DeviceBuffer(s[0] ** 3 * 4)
        self.g = {k: Tensor(t.n, t.ld) for k, t in self._acts().items()}
        self.gpred = Tensor(s[0], 1)
        ws = 0
        for l in L:
            ws = max(ws, self.lib.lgdhip_conv3d_wgrad_workspace_bytes(l.n, l.cin, l.ksize, l.cout, 0),
                     # the forward convolution's and the input gradient's (cin and cout swapped) split partials
                     self.lib.lgdhip_conv3d_forward_workspace_bytes(l.n, l.cin, l.ksize, l.cout),
                     self.lib.lgdhip_conv3d_forward_workspace_bytes(l.n, l.cout, l.ksize, l.cin))
        self.ws_bytes = ws
        self.ws = lgdhip.DeviceBuffer(ws)
        self.loss_partial = lgdhip.DeviceBuffer(self.lib.lgdhip_mse_partials(s[0] ** 3) * 4)

    # ------------------------------------------------------------------
    def _acts(self):
        return {"a0": self.a0, "cat0": self.cat0, "p1": self.p1, "a1": self.a1, "cat1": self.cat1,
                "p2": self.p2, "a2": self.a2, "cat2": self.cat2, "p3": self.p3, "a3": self.a3, "b3": self.b3,
                "a2d": self.a2d, "b2d": self.b2d, "a1d": self.a1d, "b1d": self.b1d, "a0d": self.a0d,
                "b0d": self.b0d}

    def _set_weights(self, l: Conv, w: np.ndarray, b: np.ndarray):
        wl = conv_weights_layout(np.asarray(w, dtype=np.float32), l.kp)
        if l.w32 is None:
            l.w32 = lgdhip.DeviceBuffer(wl.nbytes)
            l.wb = lgdhip.DeviceBuffer(wl.size * 2)
            l.wt = lgdhip.DeviceBuffer(l.cin * self.lib.lgdhip_conv3d_kpad(l.cout, l.ksize) * 2)
            l.bias = lgdhip.DeviceBuffer(l.cout * 4)
        l.w32.upload(wl, self.stream)
        l.wb.upload(to_bf16_bits(wl), self.stream)
        l.bias.upload(np.asarray(b, dtype=np.float32), self.stream)
        self._refresh_wt(l)

    def _refresh_wt(self, l: Conv):
        lgdhip.call("lgdhip_conv3d_dgrad_weights", l.wb.ptr, l.cin, l.cout, l.ksize, l.wt.ptr, self.stream)

    def get_weights(self, name: str):
        l = self.layers[name]
        wl = np.empty((l.cout, l.kp), np.float32)
        l.w32.download(wl, self.stream)
        taps = 27 if l.ksize == 3 else 1
        b = np.empty(l.cout, np.float32)
        l.bias.download(b, self.stream)
        return wl[:, :taps * l.cin].reshape(l.cout, taps, l.cin), b

    def set_weights(self, name: str, w, b):
        self._set_weights(self.layers[name], w, b)

    # ------------------------------------------------------------------
    def _conv(self, name, x: Tensor, xoff, y: Tensor | None, yoff, y32=None, accumulate=0):
        l = self.layers[name]
        lgdhip.call("lgdhip_conv3d_bf16_forward_ws", x.ptr, l.n, x.ld, xoff, l.cin, l.ksize, l.wb.ptr,
                    l.bias.ptr if self.use_bias else None,
                    l.cout, y.ptr if y is not None else None, y.ld if y is not None else 1, yoff,
                    y32, 1, 1 if l.relu else 0, accumulate, self.ws.ptr, self.ws_bytes, self.stream)

    def load_input(self, fields: list):
        """fields: FP32 device buffers (n^3 each) -> input tensor channels."""
        self.x.zero(self.stream)
        for c, buf in enumerate(fields):
            lgdhip.call("lgdhip_pack_bf16", buf.ptr, self.n ** 3, self.x.ptr, self.x.ld, c, self.stream)

    def forward(self):
        c0, c1, c2, c3 = self.ch
        S = self.stream
        self._conv("e0a", self.x, 0, self.a0, 0)
        self._conv("e0b", self.a0, 0, self.s0, 0)
        lgdhip.call("lgdhip_avgpool2_bf16", self.s0.ptr, self.n, c0, 0, self.p1.ptr, c0, 0, c0, S)
        self._conv("e1a", self.p1, 0, self.a1, 0)
        self._conv("e1b", self.a1, 0, self.s1, 0)
        lgdhip.call("lgdhip_avgpool2_bf16", self.s1.ptr, self.n // 2, c1, 0, self.p2.ptr, c1, 0, c1, S)
        self._conv("e2a", self.p2, 0, self.a2, 0)
        self._conv("e2b", self.a2, 0, self.s2, 0)
        lgdhip.call("lgdhip_avgpool2_bf16", self.s2.ptr, self.n // 4, c2, 0, self.p3.ptr, c2, 0, c2, S)
        self._conv("e3a", self.p3, 0, self.a3, 0)
        self._conv("e3b", self.a3, 0, self.b3, 0)
        lgdhip.call("lgdhip_upsample2_concat_bf16", self.b3.ptr, self.n // 8, c3, 0, c3, self.s2.ptr, c2, 0, c2,
                    self.cat2.ptr, self.cat2.ld, 0, S)
        self._conv("d2a", self.cat2, 0, self.a2d, 0)
        self._conv("d2b", self.a2d, 0, self.b2d, 0)
        lgdhip.call("lgdhip_upsample2_concat_bf16", self.b2d.ptr, self.n // 4, c2, 0, c2, self.s1.ptr, c1, 0, c1,
                    self.cat1.ptr, self.cat1.ld, 0, S)
        self._conv("d1a", self.cat1, 0, self.a1d, 0)
        self._conv("d1b", self.a1d, 0, self.b1d, 0)
        lgdhip.call("lgdhip_upsample2_concat_bf16", self.b1d.ptr, self.n // 2, c1, 0, c1, self.s0.ptr, c0, 0, c0,
                    self.cat0.ptr, self.cat0.ld, 0, S)
        self._conv("d0a", self.cat0, 0, self.a0d, 0)
        self._conv("d0b", self.a0d, 0, self.b0d, 0)
        self._conv("out", self.b0d, 0, None, 0, y32=self.pred.ptr)
        return self.pred

    def loss(self, target, read: bool = True) -> float:
        """MSE against an FP32 device buffer; sets the prediction gradient.
        read=False skips the host read (and its synchronisation): returns nan."""
        nv = self.n ** 3
        lgdhip.call("lgdhip_mse_loss", self.pred.ptr, target.ptr, nv, self.gpred.ptr, 1, self.loss_partial.ptr,
                    self.stream)
        if not read:
            return float("nan")
        parts = np.empty(self.lib.lgdhip_mse_partials(nv), np.float32)
        self.loss_partial.download(parts, self.stream)
        return float(parts.astype(np.float64).sum() / nv)

    def _back(self, name, x: Tensor, xoff, g_out: Tensor, goff, y_act: Tensor | None, yoff, g_in: Tensor | None,
              ginoff, accumulate_w: int, accumulate_in: int = 0):
        """Backward of one conv: ReLU mask on g_out (if relu), weight/bias
        gradient, and the input gradient into g_in (accumulated if asked)."""
        l = self.layers[name]
        S = self.stream
        if l.relu:
            lgdhip.call("lgdhip_relu_backward_bf16", g_out.ptr, g_out.ld, goff, y_act.ptr, y_act.ld, yoff, l.cout,
                        l.n ** 3, S)
        lgdhip.call("lgdhip_conv3d_bf16_wgrad", x.ptr, l.n, x.ld, xoff, l.cin, l.ksize, g_out.ptr, g_out.ld, goff,
                    l.cout, l.gw.ptr, l.gb.ptr if self.use_bias else None, accumulate_w, 0, self.ws.ptr,
                    self.ws_bytes, S)
        if g_in is not None:
            lgdhip.call("lgdhip_conv3d_bf16_forward_ws", g_out.ptr, l.n, g_out.ld, goff, l.cout, l.ksize,
                        l.wt.ptr, None, l.cin, g_in.ptr, g_in.ld, ginoff, None, 1, 0, accumulate_in, self.ws.ptr,
                        self.ws_bytes, S)

    def backward(self, accumulate: int = 0):
        """Gradients of the last loss() w.r.t. every weight (accumulate adds
        to the previous sample's gradients, for mini-batches)."""
        c0, c1, c2, c3 = self.ch
        S = self.stream
        G = self.g
        acc = accumulate
        self._back("out", self.b0d, 0, self.gpred, 0, None, 0, G["b0d"], 0, acc)
        self._back("d0b", self.a0d, 0, G["b0d"], 0, self.b0d, 0, G["a0d"], 0, acc)
        self._back("d0a", self.cat0, 0, G["a0d"], 0, self.a0d, 0, G["cat0"], 0, acc)
        lgdhip.call("lgdhip_upsample2_backward_bf16", G["cat0"].ptr, self.n // 2, G["cat0"].ld, 0, G["b1d"].ptr,
                    c1, 0, c1, 0, S)
        self._back("d1b", self.a1d, 0, G["b1d"], 0, self.b1d, 0, G["a1d"], 0, acc)
        self._back("d1a", self.cat1, 0, G["a1d"], 0, self.a1d, 0, G["cat1"], 0, acc)
        lgdhip.call("lgdhip_upsample2_backward_bf16", G["cat1"].ptr, self.n // 4, G["cat1"].ld, 0, G["b2d"].ptr,
                    c2, 0, c2, 0, S)
        self._back("d2b", self.a2d, 0, G["b2d"], 0, self.b2d, 0, G["a2d"], 0, acc)
        self._back("d2a", self.cat2, 0, G["a2d"], 0, self.a2d, 0, G["cat2"], 0, acc)
        lgdhip.call("lgdhip_upsample2_backward_bf16", G["cat2"].ptr, self.n // 8, G["cat2"].ld, 0, G["b3"].ptr,
                    c3, 0, c3, 0, S)
        self._back("e3b", self.a3, 0, G["b3"], 0, self.b3, 0, G["a3"], 0, acc)
        self._back("e3a", self.p3, 0, G["a3"], 0, self.a3, 0, G["p3"], 0, acc)
        lgdhip.call("lgdhip_avgpool2_backward_bf16", G["p3"].ptr, self.n // 4, c2, 0, G["cat2"].ptr,
                    G["cat2"].ld, c3, c2, 1, S)
        self._back("e2b", self.a2, 0, G["cat2"], c3, self.s2, 0, G["a2"], 0, acc)
        self._back("e2a", self.p2, 0, G["a2"], 0, self.a2, 0, G["p2"], 0, acc)
        lgdhip.call("lgdhip_avgpool2_backward_bf16", G["p2"].ptr, self.n // 2, c1, 0, G["cat1"].ptr,
                    G["cat1"].ld, c2, c1, 1, S)
        self._back("e1b", self.a1, 0, G["cat1"], c2, self.s1, 0, G["a1"], 0, acc)
        self._back("e1a", self.p1, 0, G["a1"], 0, self.a1, 0, G["p1"], 0, acc)
        lgdhip.call("lgdhip_avgpool2_backward_bf16", G["p1"].ptr, self.n, c0, 0, G["cat0"].ptr, G["cat0"].ld, c1,
                    c0, 1, S)
        self._back("e0b", self.a0, 0, G["cat0"], c1, self.s0, 0, G["a0"], 0, acc)
        self._back("e0a", self.x, 0, G["a0"], 0, self.a0, 0, None, 0, acc)

    def adam(self, lr: float = 1e-3, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        """One Adam step on every layer (Adam is invariant to the gradient's
        scale, so accumulated mini-batch sums need no division)."""
        self.t += 1
        S = self.stream
        for name in self.order:
            l = self.layers[name]
            taps = 27 if l.ksize == 3 else 1
            lgdhip.call("lgdhip_adam_step", l.w32.ptr, l.m[0].ptr, l.v[0].ptr, l.gw.ptr, l.cout * l.kp,
                        taps * l.cin, l.kp, lr, b1, b2, eps, self.t, l.wb.ptr, S)
            if self.use_bias:
                lgdhip.call("lgdhip_adam_step", l.bias.ptr, l.m[1].ptr, l.v[1].ptr, l.gb.ptr, l.cout, 0, 0, lr, b1,
                            b2, eps, self.t, None, S)
            self._refresh_wt(l)

    def grads(self, name: str):
        l = self.layers[name]
        g = np.empty((l.cout, l.kp), np.float32)
        l.gw.download(g, self.stream)
        gb = np.empty(l.cout, np.float32)
        l.gb.download(gb, self.stream)
        taps = 27 if l.ksize == 3 else 1
        return g[:, :taps * l.cin].reshape(l.cout, taps, l.cin), gb

    def flops_forward(self) -> float:
        f = 0.0
        for name in self.order:
            l = self.layers[name]
            taps = 27 if l.ksize == 3 else 1
            f += 2.0 * taps * l.cin * l.cout * l.n ** 3
        return f
