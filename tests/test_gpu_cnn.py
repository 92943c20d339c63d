"""Deblurring network kernels (rtm_cnn.hip) against a float64 CPU reference
(torch) that rounds to BF16 where the kernels do: the inputs, the weights,
each hidden activation and each back-propagated output gradient.  The GPU
accumulates BF16 products in FP32 on the matrix cores, the reference in
float64, so the one-layer network agrees to FP32 rounding; deeper networks
may see a rare 1-ulp BF16 flip of an intermediate value, hence the looser
bound there."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

pytestmark = pytest.mark.gpu


def bf(a):
    """Round to BF16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def unpack(p, ch):
    out, o = [], 0
    for i in range(len(ch) - 1):
        ci, co = ch[i], ch[i + 1]
        W = p[o:o + co * ci * 9].reshape(co, ci, 3, 3)
        o += co * ci * 9
        b = p[o:o + co]
        o += co
        out.append((W.astype(np.float64), b.astype(np.float64)))
    return out


def reference(p, ch, x, t=None, dil=None):
    """Forward (and with targets, MSE loss and gradient) with BF16 rounding."""
    layers = unpack(p, ch)
    dil = dil or [1] * len(layers)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a))  # noqa: E731
    acts = [bf(x)]
    for i, (W, b) in enumerate(layers):
        z = F.conv2d(T(acts[-1]), T(bf(W)), T(b), padding=dil[i], dilation=dil[i]).numpy()
        if i + 1 < len(layers):
            acts.append(bf(np.maximum(z, 0.0)))
        else:
            y = z[:, 0] + x[:, 0].astype(np.float64)
    if t is None:
        return y
    n = y.size
    loss = float(np.mean((y - t) ** 2))
    grads = [None] * len(layers)
    dz = bf((2.0 / n) * (y - t))[:, None]
    for i in range(len(layers) - 1, -1, -1):
        W, b = layers[i]
        gw = torch.nn.grad.conv2d_weight(T(acts[i]), W.shape, T(dz), padding=dil[i],
                                         dilation=dil[i]).numpy()
        grads[i] = np.concatenate([gw.ravel(), dz.sum(axis=(0, 2, 3))])
        if i > 0:
            dx = torch.nn.grad.conv2d_input(acts[i].shape, T(bf(W)), T(dz), padding=dil[i],
                                            dilation=dil[i]).numpy()
            dz = bf(dx * (acts[i] > 0))
    return y, loss, np.concatenate(grads)


def _net(ch, seed=0, npix=4096 * 4, rows=256, dil=None):
    from rtm_gpu_amd import cnn
    net = cnn.Deblur(ch, npix, rows, dilation=dil)
    p = net.init_params(seed=seed, last_zero=False)
    p[:] += np.random.default_rng(seed + 1).standard_normal(p.size).astype(np.float32) * 0.01
    net.set_params(p)
    return net, p


def _data(n, c, h, w, seed=3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, c, h, w)).astype(np.float32), \
        rng.standard_normal((n, h, w)).astype(np.float32)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("ch", [[2, 1], [1, 1], [3, 1]])
def test_one_layer_forward_and_gradient(ch):
    net, p = _net(ch)
    x, t = _data(3, ch[0], 20, 37)
    y = net.forward(x)
    yr, lr, gr = reference(p, ch, x, t)
    assert _rel(y, yr) < 1e-6
    loss, g = net.loss_grad(x, t)
    assert abs(loss - lr) <= 1e-6 * lr
    assert _rel(g, gr) < 1e-4


@pytest.mark.parametrize("ch", [[2, 16, 16, 16, 1], [2, 8, 12, 1], [3, 32, 32, 32, 1],
                                [2, 20, 32, 12, 1]])
def test_deep_forward_and_gradient(ch):
    net, p = _net(ch, seed=5)
    x, t = _data(2, ch[0], 24, 50, seed=4)
    y = net.forward(x)
    yr, lr, gr = reference(p, ch, x, t)
    print(f"{ch}: forward rel {_rel(y, yr):.2e}", flush=True)
    assert _rel(y, yr) < 1e-4
    loss, g = net.loss_grad(x, t)
    print(f"{ch}: loss {loss:.6e} / {lr:.6e}, gradient rel {_rel(g, gr):.2e}")
    assert abs(loss - lr) <= 1e-4 * lr
    assert _rel(g, gr) < 5e-3
    # every layer's block of the gradient, not only the largest one
    o = 0
    for i in range(len(ch) - 1):
        k = ch[i + 1] * ch[i] * 9 + ch[i + 1]
        assert _rel(g[o:o + k], gr[o:o + k]) < 1e-2, i
        o += k


@pytest.mark.parametrize("dil", [[1, 2, 4, 1], [3, 1, 8, 2]])
def test_dilated_forward_and_gradient(dil):
    ch = [3, 16, 8, 16, 1]
    net, p = _net(ch, seed=6, dil=dil)
    x, t = _data(2, ch[0], 30, 45, seed=5)
    y = net.forward(x)
    yr, lr, gr = reference(p, ch, x, t, dil)
    print(f"dilation {dil}: forward rel {_rel(y, yr):.2e}", flush=True)
    assert _rel(y, yr) < 1e-4
    loss, g = net.loss_grad(x, t)
    print(f"dilation {dil}: gradient rel {_rel(g, gr):.2e}")
    assert abs(loss - lr) <= 1e-4 * lr
    assert _rel(g, gr) < 5e-3


def test_adam_step():
    ch = [2, 8, 1]
    net, p = _net(ch, seed=2)
    x, t = _data(2, 2, 16, 16, seed=8)
    _, g = net.loss_grad(x, t)
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    net.adam_step(lr, b1, b2, eps)
    p1 = net.params()
    m = (1 - b1) * g.astype(np.float64)
    v = (1 - b2) * g.astype(np.float64) ** 2
    exp = p - lr * (m / (1 - b1)) / (np.sqrt(v / (1 - b2)) + eps)
    assert np.allclose(p1, exp, rtol=1e-5, atol=1e-7)


def test_save_and_load(tmp_path):
    from rtm_gpu_amd import cnn
    net, p = _net([3, 32, 8, 1], seed=4, dil=[2, 1, 3])
    x, _ = _data(2, 3, 18, 33, seed=2)
    y = net.forward(x)
    net.save(tmp_path / "net.npz")
    back = cnn.Deblur.load(tmp_path / "net.npz", 4096, 64)
    assert back.channels == [3, 32, 8, 1] and back.dilation == [2, 1, 3]
    assert np.array_equal(back.params(), p)
    assert np.array_equal(back.forward(x), y)


def test_zero_network_is_the_identity_on_channel_zero():
    from rtm_gpu_amd import cnn
    net = cnn.Deblur([2, 16, 1], 4096, 64)
    x, _ = _data(2, 2, 12, 30)
    assert np.array_equal(net.forward(x), x[:, 0])
    net.init_params(seed=1)  # last layer zero: still the identity
    assert np.array_equal(net.forward(x), x[:, 0])


def test_fit_learns_a_deconvolution():
    """Targets are sparse spikes, inputs the spikes blurred by a fixed 3x3
    kernel (and a depth channel): training must cut the loss well below the
    identity's."""
    from scipy.ndimage import convolve

    from rtm_gpu_amd import cnn
    rng = np.random.default_rng(0)
    n, h, w = 64, 24, 40
    t = (rng.random((n, h, w)) < 0.05) * rng.standard_normal((n, h, w))
    k = np.array([[0.05, 0.1, 0.05], [0.1, 0.4, 0.1], [0.05, 0.1, 0.05]])
    blur = np.stack([convolve(ti, k, mode="constant") for ti in t])
    x = np.stack([blur, np.broadcast_to(np.linspace(0, 1, h)[:, None], (n, h, w))], 1)
    x, t = x.astype(np.float32), t.astype(np.float32)
    net = cnn.Deblur([2, 16, 16, 1], 16 * h * w, 16 * h)
    net.init_params(seed=3)
    ident = float(np.mean((x[:, 0] - t) ** 2))
    hist = net.fit(x, t, epochs=60, batch=16, lr=3e-3, seed=1)
    print(f"identity loss {ident:.4e}, epoch 1 {hist[0]:.4e}, last {hist[-1]:.4e}")
    assert hist[-1] < 0.3 * ident
    y = net.forward(x[:8])
    assert float(np.mean((y - t[:8]) ** 2)) < 0.4 * ident
