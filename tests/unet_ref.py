"""FP64 reference U-Net (test infrastructure): the same architecture as
libgeodecomp_amd/unet.py, built from the oracle's FP64 convolutions and
numpy pooling/upsampling, used to check the GPU network's forward values and
weight gradients.

With bf16=True the reference rounds to BF16 (nearest even) exactly where the
GPU network stores BF16 tensors -- every activation, every pooled value and
every activation gradient -- and computes everything else in FP64, so that
it differs from the GPU only by FP32 accumulation order; with bf16=False it
is the plain FP64 network (the GPU result then differs by BF16 precision)."""
import numpy as np

import oracle_lib
from libgeodecomp_amd.unet import bf16_round


def avgpool(x):
    m = x.shape[0] // 2
    return x.reshape(m, 2, m, 2, m, 2, -1).mean(axis=(1, 3, 5))


def avgpool_back(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def upsample(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def upsample_back(g):
    m = g.shape[0] // 2
    return g.reshape(m, 2, m, 2, m, 2, -1).sum(axis=(1, 3, 5))


class RefUNet:
    def __init__(self, weights, bf16=False):
        """weights: name -> (w[cout,taps,cin], b[cout]) as float64."""
        self.W = weights
        self.q = (lambda a: bf16_round(a).astype(np.float64)) if bf16 else (lambda a: a)

    def conv(self, name, x, relu=True):
        w, b = self.W[name]
        ks = 1 if w.shape[1] == 1 else 3
        z = oracle_lib.conv3d_forward(x, w, b, ks)
        return self.q(np.maximum(z, 0.0)) if relu else z

    def forward(self, x):
        q = self.q
        A = {"x": x}
        A["a0"] = self.conv("e0a", x)
        A["e0b"] = self.conv("e0b", A["a0"])
        A["p1"] = q(avgpool(A["e0b"]))
        A["a1"] = self.conv("e1a", A["p1"])
        A["e1b"] = self.conv("e1b", A["a1"])
        A["p2"] = q(avgpool(A["e1b"]))
        A["a2"] = self.conv("e2a", A["p2"])
        A["e2b"] = self.conv("e2b", A["a2"])
        A["p3"] = q(avgpool(A["e2b"]))
        A["a3"] = self.conv("e3a", A["p3"])
        A["b3"] = self.conv("e3b", A["a3"])
        A["cat2"] = np.concatenate([upsample(A["b3"]), A["e2b"]], axis=-1)
        A["a2d"] = self.conv("d2a", A["cat2"])
        A["b2d"] = self.conv("d2b", A["a2d"])
        A["cat1"] = np.concatenate([upsample(A["b2d"]), A["e1b"]], axis=-1)
        A["a1d"] = self.conv("d1a", A["cat1"])
        A["b1d"] = self.conv("d1b", A["a1d"])
        A["cat0"] = np.concatenate([upsample(A["b1d"]), A["e0b"]], axis=-1)
        A["a0d"] = self.conv("d0a", A["cat0"])
        A["b0d"] = self.conv("d0b", A["a0d"])
        A["pred"] = self.conv("out", A["b0d"], relu=False)
        self.A = A
        return A["pred"][..., 0]

    def backward(self, gpred):
        """Weight and bias gradients for dL/dpred = gpred."""
        A, W, q = self.A, self.W, self.q
        G = {}

        def back(name, x, y, gy, relu=True, need_dx=True):
            w, _ = W[name]
            ks = 1 if w.shape[1] == 1 else 3
            gz = gy * (y > 0) if relu else gy
            G[name] = oracle_lib.conv3d_wgrad(x, gz, ks)
            return q(oracle_lib.conv3d_dgrad(gz, w, ks)) if need_dx else None

        c1, c2, c3 = A["b1d"].shape[-1], A["b2d"].shape[-1], A["b3"].shape[-1]
        g = back("out", A["b0d"], None, q(gpred[..., None]), relu=False)
        g = back("d0b", A["a0d"], A["b0d"], g)
        gcat0 = back("d0a", A["cat0"], A["a0d"], g)
        g = back("d1b", A["a1d"], A["b1d"], q(upsample_back(gcat0[..., :c1])))
        gcat1 = back("d1a", A["cat1"], A["a1d"], g)
        g = back("d2b", A["a2d"], A["b2d"], q(upsample_back(gcat1[..., :c2])))
        gcat2 = back("d2a", A["cat2"], A["a2d"], g)
        g = back("e3b", A["a3"], A["b3"], q(upsample_back(gcat2[..., :c3])))
        gp3 = back("e3a", A["p3"], A["a3"], g)
        g = back("e2b", A["a2"], A["e2b"], q(gcat2[..., c3:] + avgpool_back(gp3)))
        gp2 = back("e2a", A["p2"], A["a2"], g)
        g = back("e1b", A["a1"], A["e1b"], q(gcat1[..., c2:] + avgpool_back(gp2)))
        gp1 = back("e1a", A["p1"], A["a1"], g)
        g = back("e0b", A["a0"], A["e0b"], q(gcat0[..., c1:] + avgpool_back(gp1)))
        back("e0a", A["x"], A["a0"], g, need_dx=False)
        return G
