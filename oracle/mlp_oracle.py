"""TEST INFRASTRUCTURE ONLY: CPU oracle of the surrogate MLP
(10 -> 256 -> 256 -> 256 -> 6, SiLU) in numpy.

Imported by tests/ and bench.py's checks; never by the product package.

Two modes:
- ``emulate=True`` rounds to BF16 (round to nearest even) at exactly the
  points where the HIP kernels hold BF16 values (weights, normalised inputs,
  activations, silu' during training, output gradients, back-propagated
  gradients, tangent planes) and accumulates in float64 otherwise, so the GPU
  results differ from it only by FP32 accumulation order.
- ``emulate=False`` is the plain float64 network; it is pinned against
  finite differences and torch autograd (float64, CPU) in the tests.

The network has no reference counterpart; parity is pinned by those
mathematical checks, not by reference outputs.
"""
import numpy as np

IN, HID, OUT = 10, 256, 6
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID * HID
OFF_W4 = OFF_B3 + HID
OFF_B4 = OFF_W4 + OUT * HID
NUM_PARAMS = OFF_B4 + OUT


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(P):
    P = np.asarray(P, dtype=np.float32)
    return (P[OFF_W1:OFF_B1].reshape(HID, IN), P[OFF_B1:OFF_W2],
            P[OFF_W2:OFF_B2].reshape(HID, HID), P[OFF_B2:OFF_W3],
            P[OFF_W3:OFF_B3].reshape(HID, HID), P[OFF_B3:OFF_W4],
            P[OFF_W4:OFF_B4].reshape(OUT, HID), P[OFF_B4:NUM_PARAMS])


def pack(W1, b1, W2, b2, W3, b3, W4, b4):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3, W4.ravel(), b4])


class Norm:
    """x_n = (x - in_mean) * in_inv, y = y_n * out_std + out_mean (float32);
    a std <= 0 counts as 1; in_inv = float32(1 / float64(std))."""

    def __init__(self, in_mean, in_std, out_mean, out_std):
        f = lambda v: np.asarray(v, dtype=np.float32)
        self.in_mean = f(in_mean)
        in_std = np.where(f(in_std) > 0, f(in_std), np.float32(1))
        out_std = np.where(f(out_std) > 0, f(out_std), np.float32(1))
        self.in_inv = (1.0 / in_std.astype(np.float64)).astype(np.float32)
        self.out_mean = f(out_mean)
        self.out_std = out_std
        self.out_inv = (1.0 / out_std.astype(np.float64)).astype(np.float32)

    @staticmethod
    def identity():
        return Norm(np.zeros(IN), np.ones(IN), np.zeros(OUT), np.ones(OUT))

    def x(self, X):
        return (np.asarray(X, np.float32) - self.in_mean) * self.in_inv

    def t(self, Y):
        return (np.asarray(Y, np.float32) - self.out_mean) * self.out_inv


def silu(z):
    sg = 1.0 / (1.0 + np.exp(-z))
    return z * sg, sg * (1.0 + z * (1.0 - sg))


def _r(a, emulate):
    return bf16(a).astype(np.float64) if emulate else np.asarray(a, np.float64)


def forward(P, norm, X, emulate=True):
    """Returns (y_n, cache): normalised outputs (n x 6) and the values the
    backward pass needs."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(P)
    Ws = [_r(W, emulate) for W in (W1, W2, W3, W4)]
    bs = [np.asarray(b, np.float64) for b in (b1, b2, b3, b4)]
    xn = _r(norm.x(X), emulate)
    hs, ss = [xn], []
    a = xn
    for l in range(3):
        z = a @ Ws[l].T + bs[l]
        h, s = silu(z)
        h = _r(h, emulate)
        ss.append(_r(s, emulate))
        hs.append(h)
        a = h
    y = a @ Ws[3].T + bs[3]
    return y, dict(Ws=Ws, hs=hs, ss=ss)


def loss_and_grad(P, norm, X, Y, emulate=True):
    """Mean squared error of the normalised outputs over n x 6 entries and
    its gradient w.r.t. the packed parameters (float64)."""
    n = len(X)
    y, c = forward(P, norm, X, emulate)
    t = norm.t(Y).astype(np.float64)
    r = y - t
    loss = float(np.sum(r * r) / (6 * n))
    gscale = float(np.float32(2.0) / (np.float32(6.0) * np.float32(n))) if emulate \
        else 2.0 / (6 * n)
    d = _r(r * gscale, emulate)
    Ws, hs, ss = c["Ws"], c["hs"], c["ss"]
    gW, gb = [None] * 4, [None] * 4
    for l in (3, 2, 1, 0):
        gW[l] = d.T @ hs[l]
        gb[l] = d.sum(axis=0)
        if l > 0:
            d = _r((d @ Ws[l]) * ss[l - 1], emulate)
    return loss, pack(gW[0], gb[0], gW[1], gb[1], gW[2], gb[2], gW[3], gb[3])


def predict(P_list, norm, X, emulate=True, jacobian=False):
    """Ensemble mean of the raw outputs (n x 6), the ensemble standard
    deviation (n x 6) and, if asked, the mean Jacobian
    J[s, i, j] = d y_i / d e_j over the six strain inputs, propagated in
    forward mode with the tangent planes rounded where the kernel rounds
    them."""
    ys, Js = [], []
    for P in P_list:
        W1, b1, W2, b2, W3, b3, W4, b4 = unpack(P)
        Ws = [_r(W, emulate) for W in (W1, W2, W3, W4)]
        bs = [np.asarray(b, np.float64) for b in (b1, b2, b3, b4)]
        a = _r(norm.x(X), emulate)
        n = len(a)
        T = np.zeros((6, n, IN))
        for j in range(6):
            T[j, :, j] = 1.0
        for l in range(3):
            z = a @ Ws[l].T + bs[l]
            h, s = silu(z)
            a = _r(h, emulate)
            if jacobian:
                T = np.stack([_r((T[j] @ Ws[l].T) * s, emulate) for j in range(6)])
        yn = a @ Ws[3].T + bs[3]
        ys.append(yn)
        if jacobian:
            Js.append(np.stack([T[j] @ Ws[3].T for j in range(6)], axis=2))  # n x 6 x 6
    ys = np.array(ys)
    mean = ys.mean(axis=0)
    sd = np.sqrt(np.maximum((ys ** 2).mean(axis=0) - mean ** 2, 0))
    out_std = norm.out_std.astype(np.float64)
    Y = mean * out_std + norm.out_mean
    spread = sd * out_std
    if not jacobian:
        return Y, spread, None
    J = np.array(Js).mean(axis=0) * out_std[None, :, None] * \
        norm.in_inv[:6].astype(np.float64)[None, None, :]
    return Y, spread, J


def adam_step(P, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (float64) as the kernel applies it."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    bc1, bc2 = 1 - b1 ** t, 1 - b2 ** t
    return P - lr * (m / bc1) / (np.sqrt(v / bc2) + eps), m, v
