"""U-Net kernels: FP64 oracle pinning (CPU) and GPU parity of the MFMA
convolutions, their gradients, pooling, loss, Adam and the whole network."""
import ctypes as C

import numpy as np
import pytest

from libgeodecomp_amd import lgdhip
from libgeodecomp_amd.unet import bf16_round, conv_weights_layout, from_bf16_bits, to_bf16_bits


# ----------------------------------------------------------------------
# oracle pinning (CPU)
# ----------------------------------------------------------------------
def test_oracle_conv_delta_kernel_shifts(oracle):
    rng = np.random.default_rng(0)
    n, c = 5, 3
    x = rng.standard_normal((n, n, n, c))
    # tap (dz,dy,dx) = (1,0,-1): tap index (2)*9 + 1*3 + 0 = 21; identity over channels
    w = np.zeros((c, 27, c))
    for k in range(c):
        w[k, 21, k] = 1.0
    y = oracle.conv3d_forward(x, w)
    exp = np.zeros_like(x)
    exp[:-1, :, 1:] = x[1:, :, :-1]      # y[z,y,x] = x[z+1, y, x-1]
    assert np.array_equal(y, exp)


def test_oracle_conv_adjoint_identities(oracle):
    rng = np.random.default_rng(1)
    n, cin, cout = 4, 3, 5
    x = rng.standard_normal((n, n, n, cin))
    w = rng.standard_normal((cout, 27, cin))
    g = rng.standard_normal((n, n, n, cout))
    y = oracle.conv3d_forward(x, w)
    dx = oracle.conv3d_dgrad(g, w)
    dw, db = oracle.conv3d_wgrad(x, g)
    assert abs(np.vdot(y, g) - np.vdot(x, dx)) < 1e-10 * abs(np.vdot(y, g))
    assert abs(np.vdot(y, g) - np.vdot(w, dw)) < 1e-10 * abs(np.vdot(y, g))
    assert np.allclose(db, g.sum(axis=(0, 1, 2)))


def test_oracle_conv_finite_difference(oracle):
    rng = np.random.default_rng(2)
    n, cin, cout = 3, 2, 2
    x = rng.standard_normal((n, n, n, cin))
    w = rng.standard_normal((cout, 27, cin))
    t = rng.standard_normal((n, n, n, cout))

    def loss(wv):
        return 0.5 * np.sum((oracle.conv3d_forward(x, wv) - t) ** 2)

    dw, _ = oracle.conv3d_wgrad(x, oracle.conv3d_forward(x, w) - t)
    for idx in [(0, 13, 0), (1, 0, 1), (1, 26, 0)]:
        h = 1e-5
        wp, wm = w.copy(), w.copy()
        wp[idx] += h
        wm[idx] -= h
        fd = (loss(wp) - loss(wm)) / (2 * h)
        assert abs(fd - dw[idx]) < 1e-6 * max(1.0, abs(fd))


# ----------------------------------------------------------------------
# GPU
# ----------------------------------------------------------------------
def dev(hip, host):
    b = hip.DeviceBuffer(max(host.nbytes, 16))
    b.upload(host)
    return b


def run_conv(hip, x, w, bias, ksize, ld_out=None, off_out=0, relu=False, y_init=None):
    """x[n,n,n,cin] (bf16 values), w[cout,taps,cin] -> (y32, y_bf16 buffer slice)."""
    n, cin = x.shape[0], x.shape[-1]
    cout = w.shape[0]
    lib = hip.load()
    kp = lib.lgdhip_conv3d_kpad(cin, ksize)
    xb = dev(hip, to_bf16_bits(x))
    wb = dev(hip, to_bf16_bits(conv_weights_layout(w.astype(np.float32), kp)))
    bb = dev(hip, bias.astype(np.float32))
    ld = ld_out or cout
    y = dev(hip, to_bf16_bits(y_init) if y_init is not None else np.zeros(n ** 3 * ld, np.uint16))
    y32 = hip.DeviceBuffer(n ** 3 * cout * 4)
    hip.call("lgdhip_conv3d_bf16_forward", xb.ptr, n, cin, 0, cin, ksize, wb.ptr, bb.ptr, cout, y.ptr, ld, off_out,
             y32.ptr, cout, 1 if relu else 0, 1 if y_init is not None else 0, None)
    out32 = np.empty(n ** 3 * cout, np.float32)
    y32.download(out32)
    outb = np.empty(n ** 3 * ld, np.uint16)
    y.download(outb)
    return out32.reshape(n, n, n, cout), from_bf16_bits(outb).reshape(n, n, n, ld)


# 3x3x3 with cin % 8 == 0 below 64^3 (and any layer with more than 32
# output channels) runs the GEMM-tiled kernel, unsplit through this entry
# point; 16 output channels from 16^3 on the z-marching kernel; the LDS tile
# kernel takes 32 channels from 64^3 on and cin 2 / 3 (its element-load
# path); odd sizes leave partial blocks
CONV_CASES = [(8, 2, 16, 3), (16, 16, 16, 3), (12, 48, 16, 3), (8, 192, 64, 3), (16, 16, 1, 1), (4, 64, 128, 3),
              (8, 32, 192, 3), (6, 8, 24, 3), (32, 16, 16, 3), (33, 48, 16, 3), (20, 3, 16, 3), (16, 2, 16, 3),
              (24, 32, 32, 3), (17, 16, 24, 3), (64, 16, 32, 3), (9, 24, 72, 3)]


@pytest.mark.gpu
@pytest.mark.parametrize("n,cin,cout,ksize", CONV_CASES)
def test_conv_forward_vs_oracle(hip, oracle, n, cin, cout, ksize):
    print("[case] conv fwd", n, cin, cout, ksize, flush=True)
    rng = np.random.default_rng(n * 1000 + cin * 10 + cout)
    taps = 27 if ksize == 3 else 1
    x = bf16_round(rng.standard_normal((n, n, n, cin)).astype(np.float32))
    w = bf16_round((rng.standard_normal((cout, taps, cin)) / np.sqrt(taps * cin)).astype(np.float32))
    b = rng.standard_normal(cout).astype(np.float32)
    y32, yb = run_conv(hip, x, w, b, ksize)
    ref = oracle.conv3d_forward(x.astype(np.float64), w.astype(np.float64), b.astype(np.float64), ksize)
    # FP32 accumulation of K products: bound by K * 2^-24 * sum |terms|
    mag = oracle.conv3d_forward(np.abs(x).astype(np.float64), np.abs(w).astype(np.float64), np.abs(b), ksize)
    K = taps * cin
    assert np.all(np.abs(y32 - ref) <= (K + 2) * 2.0 ** -24 * mag + 1e-30)
    # the BF16 output is the FP32 result rounded to nearest even
    assert np.array_equal(yb, bf16_round(y32))


@pytest.mark.gpu
def test_conv_input_row_wider_than_cin(hip, oracle):
    """3 input channels in an 8-channel row whose other 5 hold junk (the
    U-Net's input tensor keeps its 3 fields in an 8-channel row): the LDS
    kernel loads whole 16-byte rows and must zero the channels past cin."""
    rng = np.random.default_rng(31)
    n, cin, ld, cout = 20, 3, 8, 16
    row = bf16_round(rng.standard_normal((n, n, n, ld)).astype(np.float32) * 100)
    x = row[..., :cin].copy()
    w = bf16_round((rng.standard_normal((cout, 27, cin)) / 9).astype(np.float32))
    lib = hip.load()
    kp = lib.lgdhip_conv3d_kpad(cin, 3)
    xb, wb = dev(hip, to_bf16_bits(row)), dev(hip, to_bf16_bits(conv_weights_layout(w, kp)))
    y32 = hip.DeviceBuffer(n ** 3 * cout * 4)
    hip.call("lgdhip_conv3d_bf16_forward", xb.ptr, n, ld, 0, cin, 3, wb.ptr, None, cout, None, cout, 0, y32.ptr,
             cout, 0, 0, None)
    out = np.empty(n ** 3 * cout, np.float32)
    y32.download(out)
    ref = oracle.conv3d_forward(x.astype(np.float64), w.astype(np.float64), None, 3)
    mag = oracle.conv3d_forward(np.abs(x).astype(np.float64), np.abs(w).astype(np.float64), None, 3)
    assert np.all(np.abs(out.reshape(ref.shape) - ref) <= (27 * cin + 2) * 2.0 ** -24 * mag + 1e-30)


@pytest.mark.gpu
@pytest.mark.parametrize("n,cin,cout,ld,off,xld", [(17, 16, 16, 16, 0, 16), (19, 48, 16, 48, 32, 48),
                                                   (33, 40, 16, 16, 0, 40), (21, 3, 16, 16, 0, 8),
                                                   (18, 32, 32, 32, 0, 32), (16, 16, 32, 96, 64, 16),
                                                   (64, 16, 32, 48, 16, 16)])
def test_conv_bf16_output_alone_equals_with_fp32_copy(hip, oracle, n, cin, cout, ld, off, xld):
    """BF16 output alone (16 output channels: the z-marching kernel; 32 at
    64^3: the tile kernel's LDS-transposed 16-byte stores) is byte-for-byte
    the output written next to an FP32 copy (the tile kernel with
    per-element stores), bias and ReLU included, partial z chunks (n = 33)
    and a 3-channel input in an 8-channel row of junk (xld = 8) too; the
    slice's other channels are untouched (32 channels below 64^3: the GEMM
    kernel, the same launch either way)."""
    rng = np.random.default_rng(n + cin + cout + ld)
    row = bf16_round(rng.standard_normal((n, n, n, xld)).astype(np.float32) * 50)
    row[..., :cin] = bf16_round(rng.standard_normal((n, n, n, cin)).astype(np.float32))
    x = row[..., :cin].copy()
    w = bf16_round((rng.standard_normal((cout, 27, cin)) / 10).astype(np.float32))
    b = rng.standard_normal(cout).astype(np.float32)
    lib = hip.load()
    kp = lib.lgdhip_conv3d_kpad(cin, 3)
    xb, wb, bb = dev(hip, to_bf16_bits(row)), dev(hip, to_bf16_bits(conv_weights_layout(w, kp))), dev(hip, b)
    init = to_bf16_bits(rng.standard_normal((n, n, n, ld)).astype(np.float32))
    outs = []
    for with32 in (False, True):
        y = dev(hip, init)
        y32 = hip.DeviceBuffer(n ** 3 * cout * 4) if with32 else None
        print("[case] conv output paths", n, cin, cout, ld, off, xld, with32, flush=True)
        hip.call("lgdhip_conv3d_bf16_forward", xb.ptr, n, xld, 0, cin, 3, wb.ptr, bb.ptr, cout, y.ptr, ld, off,
                 y32.ptr if with32 else None, cout, 1, 0, None)
        host = np.empty(n ** 3 * ld, np.uint16)
        y.download(host)
        outs.append(host.reshape(n, n, n, ld))
    assert np.array_equal(outs[0], outs[1])
    other = np.ones(ld, bool)
    other[off:off + cout] = False
    assert np.array_equal(outs[0][..., other], init.reshape(n, n, n, ld)[..., other])
    ref = np.maximum(oracle.conv3d_forward(x.astype(np.float64), w.astype(np.float64), b.astype(np.float64), 3), 0)
    got = from_bf16_bits(outs[0][..., off:off + cout]).astype(np.float64)
    assert np.all(np.abs(got - ref) <= 2.0 ** -8 * np.abs(ref) + 1e-3)


@pytest.mark.gpu
@pytest.mark.parametrize("cout,ld,off", [(32, 32, 0), (48, 64, 8)])
def test_conv_zm_slices(hip, oracle, cout, ld, off):
    """Wider outputs on volumes of at least 96^3 run the z-marching kernel
    once per 16-channel slice (the BF16-only output).  32 channels: the same
    k order as the LDS tile kernel that writes the FP32 copy, so both outputs
    are byte-equal; 48 channels (the FP32 copy from the GEMM-tiled kernel)
    into a 64-channel row at channel 8: the other channels untouched.  Both
    checked against the FP64 oracle on two corner cubes (the oracle's
    convolution of the corner 10^3 of the input equals the full one on the
    inner 9^3 next to the volume's faces)."""
    n, cin = 96, 16
    rng = np.random.default_rng(cout * 7 + ld)
    x = bf16_round(rng.standard_normal((n, n, n, cin)).astype(np.float32))
    w = bf16_round((rng.standard_normal((cout, 27, cin)) / 12).astype(np.float32))
    b = rng.standard_normal(cout).astype(np.float32)
    lib = hip.load()
    kp = lib.lgdhip_conv3d_kpad(cin, 3)
    xb, wb, bb = dev(hip, to_bf16_bits(x)), dev(hip, to_bf16_bits(conv_weights_layout(w, kp))), dev(hip, b)
    init = to_bf16_bits(rng.standard_normal((n, n, n, ld)).astype(np.float32))
    outs = []
    for with32 in (False, True):
        y = dev(hip, init)
        y32 = hip.DeviceBuffer(n ** 3 * cout * 4) if with32 else None
        print("[case] conv z-march slices", n, cin, cout, ld, off, with32, flush=True)
        hip.call("lgdhip_conv3d_bf16_forward", xb.ptr, n, cin, 0, cin, 3, wb.ptr, bb.ptr, cout, y.ptr, ld, off,
                 y32.ptr if with32 else None, cout, 1, 0, None)
        host = np.empty(n ** 3 * ld, np.uint16)
        y.download(host)
        outs.append(host.reshape(n, n, n, ld))
    if cout == 32:
        assert np.array_equal(outs[0], outs[1])
    other = np.ones(ld, bool)
    other[off:off + cout] = False
    for o in outs:
        assert np.array_equal(o[..., other], init.reshape(n, n, n, ld)[..., other])
    for lo in (0, n - 10):
        c = x[lo:lo + 10, lo:lo + 10, lo:lo + 10]
        ref = np.maximum(oracle.conv3d_forward(c.astype(np.float64), w.astype(np.float64), b.astype(np.float64), 3), 0)
        keep = slice(0, 9) if lo == 0 else slice(1, 10)
        ref = ref[keep, keep, keep]
        for o in outs:
            got = from_bf16_bits(o[lo:lo + 10, lo:lo + 10, lo:lo + 10, off:off + cout])[keep, keep, keep]
            assert np.all(np.abs(got.astype(np.float64) - ref) <= 2.0 ** -8 * np.abs(ref) + 1e-3)


def run_conv_split(hip, x, w, bias, ld_out=None, off_out=0, relu=False, y_init=None):
    """run_conv through lgdhip_conv3d_bf16_forward_ws with the workspace the
    layer asks for; returns (y32, y_bf16, workspace bytes)."""
    n, cin = x.shape[0], x.shape[-1]
    cout = w.shape[0]
    lib = hip.load()
    kp = lib.lgdhip_conv3d_kpad(cin, 3)
    wsb = lib.lgdhip_conv3d_forward_workspace_bytes(n, cin, 3, cout)
    ws = hip.DeviceBuffer(max(wsb, 16))
    xb = dev(hip, to_bf16_bits(x))
    wb = dev(hip, to_bf16_bits(conv_weights_layout(w.astype(np.float32), kp)))
    bb = dev(hip, bias.astype(np.float32))
    ld = ld_out or cout
    y = dev(hip, to_bf16_bits(y_init) if y_init is not None else np.zeros(n ** 3 * ld, np.uint16))
    y32 = hip.DeviceBuffer(n ** 3 * cout * 4)
    hip.call("lgdhip_conv3d_bf16_forward_ws", xb.ptr, n, cin, 0, cin, 3, wb.ptr, bb.ptr, cout, y.ptr, ld, off_out,
             y32.ptr, cout, 1 if relu else 0, 1 if y_init is not None else 0, ws.ptr, wsb, None)
    out32 = np.empty(n ** 3 * cout, np.float32)
    y32.download(out32)
    outb = np.empty(n ** 3 * ld, np.uint16)
    y.download(outb)
    return out32.reshape(n, n, n, cout), from_bf16_bits(outb).reshape(n, n, n, ld), wsb


@pytest.mark.gpu
@pytest.mark.parametrize("n,cin,cout", [(8, 192, 64), (16, 64, 64), (16, 128, 128), (12, 48, 80)])
def test_conv_split_k_vs_oracle(hip, oracle, n, cin, cout):
    """Small volumes with a workspace: the GEMM kernel's k range split over
    work-groups (3-20 slices here) and the partials added in slice order.
    Same bound as the unsplit kernels (each output is still an FP32 sum of
    the same products, grouped by slice); the BF16 output is the FP32
    result rounded; two runs are identical (fixed-order reduction)."""
    print("[case] conv split-k", n, cin, cout, flush=True)
    rng = np.random.default_rng(n * 7 + cin + cout)
    x = bf16_round(rng.standard_normal((n, n, n, cin)).astype(np.float32))
    w = bf16_round((rng.standard_normal((cout, 27, cin)) / np.sqrt(27 * cin)).astype(np.float32))
    b = rng.standard_normal(cout).astype(np.float32)
    y32, yb, wsb = run_conv_split(hip, x, w, b)
    assert wsb > 0  # the layer splits
    ref = oracle.conv3d_forward(x.astype(np.float64), w.astype(np.float64), b.astype(np.float64), 3)
    mag = oracle.conv3d_forward(np.abs(x).astype(np.float64), np.abs(w).astype(np.float64), np.abs(b), 3)
    K = 27 * cin
    assert np.all(np.abs(y32 - ref) <= (K + 2) * 2.0 ** -24 * mag + 1e-30)
    assert np.array_equal(yb, bf16_round(y32))
    y32b, _, _ = run_conv_split(hip, x, w, b)
    assert np.array_equal(y32, y32b)
    # against the unsplit launch: FP32 rounding apart
    u32, _ = run_conv(hip, x, w, b, 3)
    assert np.all(np.abs(u32 - y32) <= 2 * (K + 2) * 2.0 ** -24 * mag + 1e-30)


@pytest.mark.gpu
def test_conv_split_k_relu_slice_accumulate(hip, oracle):
    """The split path's reduction applies ReLU, the accumulation into an
    existing BF16 slice and the FP32 copy like the unsplit epilogue, and
    leaves the slice's other channels alone."""
    rng = np.random.default_rng(44)
    n, cin, cout, ld, off = 16, 64, 48, 96, 32
    x = bf16_round(rng.standard_normal((n, n, n, cin)).astype(np.float32))
    w = bf16_round((rng.standard_normal((cout, 27, cin)) / 40).astype(np.float32))
    b = rng.standard_normal(cout).astype(np.float32)
    init = bf16_round(rng.standard_normal((n, n, n, ld)).astype(np.float32))
    y32, yb, wsb = run_conv_split(hip, x, w, b, ld_out=ld, off_out=off, relu=True, y_init=init)
    assert wsb > 0
    ref = np.maximum(oracle.conv3d_forward(x, w.astype(np.float64), b.astype(np.float64), 3), 0)
    got = y32.astype(np.float64) - init[..., off:off + cout]
    assert np.all(np.abs(got - ref) <= 1e-5 * (1 + np.abs(init[..., off:off + cout]) + ref))
    assert np.array_equal(yb[..., :off], init[..., :off])
    assert np.array_equal(yb[..., off + cout:], init[..., off + cout:])
    assert np.array_equal(yb[..., off:off + cout], bf16_round(y32))


@pytest.mark.gpu
@pytest.mark.parametrize("n", [8, 19])
def test_conv_relu_slice_accumulate(hip, oracle, n):
    rng = np.random.default_rng(9 + n)
    cin, cout, ld, off = 16, 16, 48, 32
    x = bf16_round(rng.standard_normal((n, n, n, cin)).astype(np.float32))
    w = bf16_round((rng.standard_normal((cout, 27, cin)) / 12).astype(np.float32))
    b = np.zeros(cout, np.float32)
    init = bf16_round(rng.standard_normal((n, n, n, ld)).astype(np.float32))
    y32, yb = run_conv(hip, x, w, b, 3, ld_out=ld, off_out=off, relu=True, y_init=init)
    ref = np.maximum(oracle.conv3d_forward(x, w.astype(np.float64), None, 3), 0)
    # y32 = init + relu(conv) in FP32; the BF16 slice is that value rounded
    got = y32.astype(np.float64) - init[..., off:]
    assert np.all(np.abs(got - ref) <= 1e-5 * (1 + np.abs(init[..., off:]) + ref))
    assert np.array_equal(yb[..., :off], init[..., :off])
    assert np.array_equal(yb[..., off:], bf16_round(y32))


@pytest.mark.gpu
@pytest.mark.parametrize("n,cin,cout", [(8, 16, 16), (8, 48, 16), (4, 192, 64), (8, 16, 1), (16, 16, 16),
                                        (18, 32, 16), (16, 192, 64), (12, 64, 128)])
def test_conv_dgrad_vs_oracle(hip, oracle, n, cin, cout):
    print("[case] conv dgrad", n, cin, cout, flush=True)
    rng = np.random.default_rng(cin + cout)
    ksize = 3 if cout > 1 else 1
    taps = 27 if ksize == 3 else 1
    w = bf16_round((rng.standard_normal((cout, taps, cin)) / 8).astype(np.float32))
    g = bf16_round(rng.standard_normal((n, n, n, cout)).astype(np.float32))
    lib = hip.load()
    kp, kpt = lib.lgdhip_conv3d_kpad(cin, ksize), lib.lgdhip_conv3d_kpad(cout, ksize)
    wb = dev(hip, to_bf16_bits(conv_weights_layout(w, kp)))
    wt = hip.DeviceBuffer(cin * kpt * 2)
    hip.call("lgdhip_conv3d_dgrad_weights", wb.ptr, cin, cout, ksize, wt.ptr, None)
    gb = dev(hip, to_bf16_bits(g))
    dx32 = hip.DeviceBuffer(n ** 3 * cin * 4)
    hip.call("lgdhip_conv3d_bf16_forward", gb.ptr, n, cout, 0, cout, ksize, wt.ptr, None, cin, None, cin, 0,
             dx32.ptr, cin, 0, 0, None)
    out = np.empty(n ** 3 * cin, np.float32)
    dx32.download(out)
    ref = oracle.conv3d_dgrad(g, w.astype(np.float64), ksize)
    mag = oracle.conv3d_dgrad(np.abs(g), np.abs(w).astype(np.float64), ksize)
    assert np.all(np.abs(out.reshape(ref.shape) - ref) <= (taps * cout + 2) * 2.0 ** -24 * mag + 1e-30)


@pytest.mark.gpu
@pytest.mark.parametrize("n,cin,cout,ksize", [(8, 16, 16, 3), (12, 2, 16, 3), (4, 192, 64, 3), (16, 16, 1, 1),
                                              (16, 32, 64, 3), (8, 64, 32, 3), (16, 48, 16, 3), (24, 16, 48, 3),
                                              (8, 192, 128, 3), (8, 32, 16, 3), (8, 64, 16, 3),
                                              (8, 16, 32, 3), (8, 32, 32, 3), (8, 16, 64, 3)])
def test_conv_wgrad_vs_oracle(hip, oracle, n, cin, cout, ksize):
    """Element-gather kernel (odd channel counts, 1x1x1, n % 8 != 0) and the
    voxel-blocked kernel (3x3x3, channels multiples of 16, n % 8 == 0, every
    tile grouping CT x IT in {1, 2, 4}^2 reached by some case)."""
    run_wgrad_case(hip, oracle, n, cin, cout, ksize, 256)


@pytest.mark.gpu
@pytest.mark.parametrize("n,cin,cout,ksize", [(12, 2, 16, 3), (16, 32, 64, 3), (8, 192, 128, 3), (24, 16, 48, 3)])
def test_conv_wgrad_default_chunk(hip, oracle, n, cin, cout, ksize):
    """chunk_vox = 0: the library's chunk for the shape (128-voxel chunks
    here), in the workspace query and the call alike."""
    run_wgrad_case(hip, oracle, n, cin, cout, ksize, 0)


def test_conv_wgrad_default_chunk_sizes():
    """The library's chunk (chunk_vox = 0) for the U-Net's levels, read off
    the workspace size: 4096 voxels while that gives 2048 waves (128^3, 48
    -> 16 channels: 512 chunks x 9 waves), smaller elsewhere (128^3 16 ->
    16, one wave per dz: 2048; 32^3 64 -> 64: 256; 16^3 128 -> 128: 128),
    never below 128."""
    lib = lgdhip.load()
    for n, cin, cout, chunk in [(128, 16, 16, 2048), (128, 48, 16, 4096), (64, 32, 32, 1024), (32, 64, 64, 256),
                                (16, 128, 128, 128), (128, 3, 16, 4096), (8, 16, 16, 128)]:
        assert (lib.lgdhip_conv3d_wgrad_workspace_bytes(n, cin, 3, cout, 0) ==
                lib.lgdhip_conv3d_wgrad_workspace_bytes(n, cin, 3, cout, chunk)), (n, cin, cout, chunk)


def run_wgrad_case(hip, oracle, n, cin, cout, ksize, chunk):
    print("[case] conv wgrad", n, cin, cout, ksize, "chunk", chunk, flush=True)
    rng = np.random.default_rng(7 * cin + cout)
    taps = 27 if ksize == 3 else 1
    x = bf16_round(rng.standard_normal((n, n, n, cin)).astype(np.float32))
    g = bf16_round(rng.standard_normal((n, n, n, cout)).astype(np.float32))
    lib = hip.load()
    kp = lib.lgdhip_conv3d_kpad(cin, ksize)
    wsb = lib.lgdhip_conv3d_wgrad_workspace_bytes(n, cin, ksize, cout, chunk)
    ws = hip.DeviceBuffer(wsb)
    xb, gb = dev(hip, to_bf16_bits(x)), dev(hip, to_bf16_bits(g))
    dw = hip.DeviceBuffer(cout * kp * 4)
    db = hip.DeviceBuffer(cout * 4)
    hip.call("lgdhip_conv3d_bf16_wgrad", xb.ptr, n, cin, 0, cin, ksize, gb.ptr, cout, 0, cout, dw.ptr, db.ptr, 0,
             chunk, ws.ptr, wsb, None)
    out = np.empty((cout, kp), np.float32)
    dw.download(out)
    outb = np.empty(cout, np.float32)
    db.download(outb)
    ref, refb = oracle.conv3d_wgrad(x, g, ksize)
    mag, _ = oracle.conv3d_wgrad(np.abs(x), np.abs(g), ksize)
    got = out[:, :taps * cin].reshape(ref.shape)
    assert np.all(out[:, taps * cin:] == 0)
    assert np.all(np.abs(got - ref) <= (n ** 3 + 64) * 2.0 ** -24 * mag + 1e-30)
    assert np.allclose(outb, refb, rtol=1e-5, atol=1e-3)


@pytest.mark.gpu
def test_conv_wgrad_strided_tensors(hip, oracle):
    """x and dy as channel ranges of wider (concatenated) buffers: ld > c,
    off > 0, as the U-Net's skip concatenations pass them."""
    n, cin, cout, ldx, offx, ldy, offy = 16, 16, 32, 48, 32, 40, 8
    rng = np.random.default_rng(3)
    x = bf16_round(rng.standard_normal((n, n, n, cin)).astype(np.float32))
    g = bf16_round(rng.standard_normal((n, n, n, cout)).astype(np.float32))
    xs = np.zeros((n, n, n, ldx), np.float32)
    xs[..., offx:offx + cin] = x
    gs = np.zeros((n, n, n, ldy), np.float32)
    gs[..., offy:offy + cout] = g
    lib = hip.load()
    kp = lib.lgdhip_conv3d_kpad(cin, 3)
    wsb = lib.lgdhip_conv3d_wgrad_workspace_bytes(n, cin, 3, cout, 512)
    ws = hip.DeviceBuffer(wsb)
    xb, gb = dev(hip, to_bf16_bits(xs)), dev(hip, to_bf16_bits(gs))
    dw = hip.DeviceBuffer(cout * kp * 4)
    db = hip.DeviceBuffer(cout * 4)
    hip.call("lgdhip_conv3d_bf16_wgrad", xb.ptr, n, ldx, offx, cin, 3, gb.ptr, ldy, offy, cout, dw.ptr, db.ptr, 0,
             512, ws.ptr, wsb, None)
    out = np.empty((cout, kp), np.float32)
    dw.download(out)
    outb = np.empty(cout, np.float32)
    db.download(outb)
    ref, refb = oracle.conv3d_wgrad(x, g, 3)
    mag, _ = oracle.conv3d_wgrad(np.abs(x), np.abs(g), 3)
    assert np.all(np.abs(out[:, :27 * cin].reshape(ref.shape) - ref) <= (n ** 3 + 64) * 2.0 ** -24 * mag + 1e-30)
    assert np.all(out[:, 27 * cin:] == 0)
    assert np.allclose(outb, refb, rtol=1e-5, atol=1e-3)


def children_sum(a, m, c):
    """FP32 sum of the 8 children of each parent voxel in the kernels' order
    (children z-major, then y, then x) of an (2m)^3 x c array."""
    blocks = a.reshape(m, 2, m, 2, m, 2, c)
    s = np.zeros((m, m, m, c), np.float32)
    for d in range(8):
        s = s + blocks[:, d >> 2, :, (d >> 1) & 1, :, d & 1, :]
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("c,pad,off", [(16, 0, 0), (24, 8, 8), (3, 2, 1), (16, 3, 0)])
def test_pool_upsample_forward_backward(hip, c, pad, off):
    """2x2x2 average pool, nearest upsample and both gradients, in channel
    slices of wider rows (row stride c + pad, channel offset off): the
    8-channel path (c, stride and offset multiples of 8) and the per-element
    path give the numpy restatement of the kernels' FP32 order bit for bit,
    with and without accumulation, and leave the other channels untouched."""
    rng = np.random.default_rng(c * 10 + pad)
    n, m = 8, 4
    ld = c + pad

    def wide(shape_vox):
        return bf16_round(rng.standard_normal(shape_vox + (ld,)).astype(np.float32))

    def run(name, src_w, dst_w, *args):
        sb, db = dev(hip, to_bf16_bits(src_w)), dev(hip, to_bf16_bits(dst_w))
        hip.call(name, sb.ptr, *args[:1], ld, off, db.ptr, ld, off, c, *args[1:], None)
        out = np.empty(dst_w.size, np.uint16)
        db.download(out)
        return from_bf16_bits(out).reshape(dst_w.shape)

    sl = slice(off, off + c)
    x = wide((n, n, n))
    y0 = wide((m, m, m))
    print(f"[case] pool/upsample c {c} ld {ld} off {off}", flush=True)
    y = run("lgdhip_avgpool2_bf16", x, y0, n)
    exp = y0.copy()
    exp[..., sl] = bf16_round(children_sum(x[..., sl], m, c) * np.float32(0.125))
    assert np.array_equal(y, exp)
    u0 = wide((n, n, n))
    u = run("lgdhip_upsample2_bf16", y, u0, m)
    exp = u0.copy()
    exp[..., sl] = np.repeat(np.repeat(np.repeat(y[..., sl], 2, 0), 2, 1), 2, 2)
    assert np.array_equal(u, exp)
    for acc in (0, 1):
        gy = wide((m, m, m))
        gx0 = wide((n, n, n))
        gx = run("lgdhip_avgpool2_backward_bf16", gy, gx0, n, acc)
        g = np.repeat(np.repeat(np.repeat(gy[..., sl], 2, 0), 2, 1), 2, 2) * np.float32(0.125)
        exp = gx0.copy()
        exp[..., sl] = bf16_round(g + gx0[..., sl] if acc else g)
        assert np.array_equal(gx, exp), ("avgpool backward", acc)
        gu = wide((n, n, n))
        gp0 = wide((m, m, m))
        gp = run("lgdhip_upsample2_backward_bf16", gu, gp0, m, acc)
        s = children_sum(gu[..., sl], m, c)
        exp = gp0.copy()
        exp[..., sl] = bf16_round(s + gp0[..., sl] if acc else s)
        assert np.array_equal(gp, exp), ("upsample backward", acc)


@pytest.mark.gpu
@pytest.mark.parametrize("cu,cs,ldx,offx,lds,offs,ldy,offy", [(32, 16, 32, 0, 16, 0, 48, 0),
                                                             (16, 8, 24, 8, 16, 8, 40, 8),
                                                             (4, 3, 5, 1, 3, 0, 9, 2)])
def test_upsample_concat(hip, cu, cs, ldx, offx, lds, offs, ldy, offy):
    """Upsampled channels and a skip tensor written as whole concatenation
    rows: y[v][offy:offy+cu] = x[parent(v)], y[v][offy+cu:offy+cu+cs] =
    skip[v], byte for byte, the row's other channels untouched (the
    8-channel path, and the per-element path for counts and offsets that are
    not multiples of 8; odd m)."""
    rng = np.random.default_rng(cu * 7 + cs + ldy)
    for m in (4, 3):
        n = 2 * m
        x = bf16_round(rng.standard_normal((m, m, m, ldx)).astype(np.float32))
        sk = bf16_round(rng.standard_normal((n, n, n, lds)).astype(np.float32))
        y0 = bf16_round(rng.standard_normal((n, n, n, ldy)).astype(np.float32))
        xb, sb, yb = dev(hip, to_bf16_bits(x)), dev(hip, to_bf16_bits(sk)), dev(hip, to_bf16_bits(y0))
        print(f"[case] upsample concat m {m} cu {cu} cs {cs} ldy {ldy} offy {offy}", flush=True)
        hip.call("lgdhip_upsample2_concat_bf16", xb.ptr, m, ldx, offx, cu, sb.ptr, lds, offs, cs, yb.ptr, ldy, offy,
                 None)
        out = np.empty(y0.size, np.uint16)
        yb.download(out)
        exp = y0.copy()
        up = np.repeat(np.repeat(np.repeat(x[..., offx:offx + cu], 2, 0), 2, 1), 2, 2)
        exp[..., offy:offy + cu] = up
        exp[..., offy + cu:offy + cu + cs] = sk[..., offs:offs + cs]
        assert np.array_equal(from_bf16_bits(out).reshape(y0.shape), exp)


@pytest.mark.gpu
def test_relu_mse_adam(hip):
    rng = np.random.default_rng(4)
    n, c = 8, 16
    x = bf16_round(rng.standard_normal((n, n, n, c)).astype(np.float32))
    xb = dev(hip, to_bf16_bits(x))
    # relu backward: zero where y <= 0
    g = bf16_round(rng.standard_normal((n, n, n, c)).astype(np.float32))
    gb = dev(hip, to_bf16_bits(g))
    hip.call("lgdhip_relu_backward_bf16", gb.ptr, c, 0, xb.ptr, c, 0, c, n ** 3, None)
    gg = np.empty(n ** 3 * c, np.uint16)
    gb.download(gg)
    assert np.array_equal(from_bf16_bits(gg).reshape(g.shape), np.where(x > 0, g, 0))
    # mse
    lib = hip.load()
    pred = rng.standard_normal(n ** 3).astype(np.float32)
    tgt = rng.standard_normal(n ** 3).astype(np.float32)
    pb, tb = dev(hip, pred), dev(hip, tgt)
    grad = hip.DeviceBuffer(n ** 3 * 2)
    parts = hip.DeviceBuffer(lib.lgdhip_mse_partials(n ** 3) * 4)
    hip.call("lgdhip_mse_loss", pb.ptr, tb.ptr, n ** 3, grad.ptr, 1, parts.ptr, None)
    ph = np.empty(lib.lgdhip_mse_partials(n ** 3), np.float32)
    parts.download(ph)
    assert abs(ph.sum() / n ** 3 - np.mean((pred - tgt) ** 2)) < 1e-5
    gh = np.empty(n ** 3, np.uint16)
    grad.download(gh)
    assert np.array_equal(from_bf16_bits(gh), bf16_round((2 * (pred - tgt) / n ** 3).astype(np.float32)))
    # adam, two steps against numpy
    w = rng.standard_normal(64).astype(np.float32)
    gw = rng.standard_normal(64).astype(np.float32)
    wd, md, vd, gd = dev(hip, w), dev(hip, np.zeros(64, np.float32)), dev(hip, np.zeros(64, np.float32)), dev(hip, gw)
    mw, vw, wr = np.zeros(64), np.zeros(64), w.astype(np.float64)
    for t in (1, 2):
        hip.call("lgdhip_adam_step", wd.ptr, md.ptr, vd.ptr, gd.ptr, 64, 0, 0, 1e-2, 0.9, 0.999, 1e-8, t, None, None)
        mw = 0.9 * mw + 0.1 * gw
        vw = 0.999 * vw + 0.001 * gw * gw
        wr = wr - 1e-2 * (mw / (1 - 0.9 ** t)) / (np.sqrt(vw / (1 - 0.999 ** t)) + 1e-8)
    wout = np.empty(64, np.float32)
    wd.download(wout)
    assert np.allclose(wout, wr, atol=1e-6)


@pytest.mark.gpu
def test_unet_gradients_vs_fp64_oracle(hip, oracle):
    """Whole network at 16^3: forward values and every layer's weight and
    bias gradient against the FP64 oracle network with the same (BF16)
    weights.  Against the reference that rounds to BF16 where the GPU stores
    BF16, the only difference left is FP32 accumulation order (K up to 5184
    products: ~3e-4 relative, enough to move a BF16 rounding of many
    elements by one unit), so the bound is 3e-2 per layer with cosine >
    0.9995 (an index or tap mistake gives O(1));
    against the plain FP64 network the BF16 activation and gradient storage
    shows as a few per cent, bounded at 25 % with cosine > 0.98."""
    import unet_ref
    from libgeodecomp_amd.unet import UNet3D

    n = 16
    rng = np.random.default_rng(11)
    net = UNet3D(n=n, seed=3)
    W = {}
    for name in net.order:
        w, b = net.get_weights(name)
        w = bf16_round(w)
        b = rng.standard_normal(b.shape).astype(np.float32) * 0.05
        net.set_weights(name, w, b)
        W[name] = (w.astype(np.float64), b.astype(np.float64))
    f = rng.standard_normal((n, n, n)).astype(np.float32)
    bnd = rng.standard_normal((n, n, n)).astype(np.float32)
    tgt = rng.standard_normal((n, n, n)).astype(np.float32)
    fb, bb, tb = dev(hip, f), dev(hip, bnd), dev(hip, tgt)
    net.load_input([fb, bb])
    pred_dev = net.forward()
    loss = net.loss(tb)
    net.backward()
    pred = np.empty(n ** 3, np.float32)
    pred_dev.download(pred)
    pred = pred.reshape(n, n, n)
    x = np.stack([bf16_round(f), bf16_round(bnd)], axis=-1).astype(np.float64)

    failures = []
    for bf16, tol_w, tol_cos, tol_pred in ((True, 3e-2, 0.9995, 5e-3), (False, 0.25, 0.98, 3e-2)):
        ref = unet_ref.RefUNet(W, bf16=bf16)
        pr = ref.forward(x)
        rel_pred = np.linalg.norm(pred - pr) / np.linalg.norm(pr)
        loss_ref = np.mean((pr - tgt) ** 2)
        G = ref.backward(2 * (pr - tgt) / n ** 3)
        print(f"[unet] bf16-emulating={bf16}: forward rel err {rel_pred:.3e}, loss gpu {loss:.6e} "
              f"ref {loss_ref:.6e}", flush=True)
        if not (rel_pred < tol_pred and abs(loss - loss_ref) < tol_pred * loss_ref):
            failures.append((bf16, "forward"))
        for name in net.order:
            gw, gb = net.grads(name)
            rw, rb = G[name]
            ew = np.linalg.norm(gw - rw) / max(np.linalg.norm(rw), 1e-30)
            eb = np.linalg.norm(gb - rb) / max(np.linalg.norm(rb), 1e-30)
            cos = np.vdot(gw, rw) / (np.linalg.norm(gw) * np.linalg.norm(rw) + 1e-300)
            print(f"  {name}: |dW err| {ew:.3e} cos {cos:.6f} |db err| {eb:.3e}", flush=True)
            if not (ew < tol_w and cos > tol_cos and eb < tol_w):
                failures.append((bf16, name))
    assert not failures, failures
