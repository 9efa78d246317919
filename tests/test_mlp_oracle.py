"""Pins the numpy MLP oracle (oracle/mlp_oracle.py) that the GPU training and
inference kernels are checked against: BF16 rounding known answers, float64
gradients against finite differences and torch autograd, the forward-mode
Jacobian against finite differences and autograd, and the Adam update's
closed form at step 1.  The network has no reference counterpart, so these
mathematical checks are what fixes it."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo


def random_params(seed):
    rng = np.random.default_rng(seed)
    fan = {mo.OFF_W1: mo.IN, mo.OFF_B1: mo.IN, mo.OFF_W2: mo.HID, mo.OFF_B2: mo.HID,
           mo.OFF_W3: mo.HID, mo.OFF_B3: mo.HID, mo.OFF_W4: mo.HID, mo.OFF_B4: mo.HID}
    bounds = sorted(fan) + [mo.NUM_PARAMS]
    P = np.empty(mo.NUM_PARAMS, np.float32)
    for a, b in zip(bounds[:-1], bounds[1:]):
        P[a:b] = rng.uniform(-1, 1, b - a) / np.sqrt(fan[a])
    return P


def random_norm(seed):
    rng = np.random.default_rng(seed)
    return mo.Norm(rng.normal(0, 0.1, mo.IN), rng.uniform(0.5, 2, mo.IN),
                   rng.normal(0, 0.1, mo.OUT), rng.uniform(0.5, 2, mo.OUT))


def test_bf16_round_to_nearest_even():
    e = np.float32(2.0 ** -8)
    vals = np.array([1.0, 1 + e, 1 + 3 * e, 1 + e + 2.0 ** -20, -1 - e, 3.0e38, 1e-30],
                    np.float32)
    got = mo.bf16(vals)
    want = np.array([1.0, 1.0, 1 + 4 * e, 1 + 2 * e, -1.0, 2.9975e38, 1e-30], np.float32)
    assert got[0] == 1.0 and got[1] == 1.0          # tie -> even (down)
    assert got[2] == np.float32(1 + 4 * e)          # tie -> even (up)
    assert got[3] == np.float32(1 + 2 * e)          # above the tie -> up
    assert got[4] == -1.0
    assert abs(got[5] / want[5] - 1) < 4e-3
    assert abs(got[6] / 1e-30 - 1) < 4e-3
    # every result is a bfloat16 (low 16 bits clear)
    assert np.all(got.view(np.uint32) & 0xFFFF == 0)


def test_gradient_float64_matches_finite_differences():
    P = random_params(1)
    nm = random_norm(2)
    rng = np.random.default_rng(3)
    X = rng.normal(0, 1, (64, mo.IN))
    Y = rng.normal(0, 1, (64, mo.OUT))
    _, g = mo.loss_and_grad(P, nm, X, Y, emulate=False)
    P64 = P.astype(np.float64)

    def loss_at(Pv):
        # float64 forward without rounding (unpack keeps float32, so use the
        # float64 path by hand)
        W = [Pv[mo.OFF_W1:mo.OFF_B1].reshape(256, 10), Pv[mo.OFF_W2:mo.OFF_B2].reshape(256, 256),
             Pv[mo.OFF_W3:mo.OFF_B3].reshape(256, 256), Pv[mo.OFF_W4:mo.OFF_B4].reshape(6, 256)]
        b = [Pv[mo.OFF_B1:mo.OFF_W2], Pv[mo.OFF_B2:mo.OFF_W3], Pv[mo.OFF_B3:mo.OFF_W4],
             Pv[mo.OFF_B4:]]
        a = nm.x(X).astype(np.float64)
        for l in range(3):
            a = mo.silu(a @ W[l].T + b[l])[0]
        y = a @ W[3].T + b[3]
        r = y - nm.t(Y).astype(np.float64)
        return np.sum(r * r) / (6 * len(X))

    rng = np.random.default_rng(4)
    picks = [mo.OFF_W1 + 7, mo.OFF_B1 + 3, mo.OFF_W2 + 1234, mo.OFF_B2 + 100,
             mo.OFF_W3 + 40000, mo.OFF_B3 + 5, mo.OFF_W4 + 777, mo.OFF_B4 + 2]
    picks += list(rng.integers(0, mo.NUM_PARAMS, 8))
    for p in picks:
        h = 1e-6
        Pp, Pm = P64.copy(), P64.copy()
        Pp[p] += h
        Pm[p] -= h
        fd = (loss_at(Pp) - loss_at(Pm)) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(abs(fd), 1e-3), (p, fd, g[p])


def test_gradient_and_jacobian_match_torch_autograd():
    torch = pytest.importorskip("torch")
    P = random_params(5)
    nm = random_norm(6)
    rng = np.random.default_rng(7)
    X = rng.normal(0, 1, (32, mo.IN))
    Y = rng.normal(0, 1, (32, mo.OUT))
    loss, g = mo.loss_and_grad(P, nm, X, Y, emulate=False)
    Pt = torch.tensor(P.astype(np.float64), requires_grad=True)

    def net(Pt, xn):
        W1 = Pt[mo.OFF_W1:mo.OFF_B1].reshape(256, 10)
        W2 = Pt[mo.OFF_W2:mo.OFF_B2].reshape(256, 256)
        W3 = Pt[mo.OFF_W3:mo.OFF_B3].reshape(256, 256)
        W4 = Pt[mo.OFF_W4:mo.OFF_B4].reshape(6, 256)
        a = torch.nn.functional.silu(xn @ W1.T + Pt[mo.OFF_B1:mo.OFF_W2])
        a = torch.nn.functional.silu(a @ W2.T + Pt[mo.OFF_B2:mo.OFF_W3])
        a = torch.nn.functional.silu(a @ W3.T + Pt[mo.OFF_B3:mo.OFF_W4])
        return a @ W4.T + Pt[mo.OFF_B4:]

    xn = torch.tensor(nm.x(X).astype(np.float64))
    t = torch.tensor(nm.t(Y).astype(np.float64))
    lt = ((net(Pt, xn) - t) ** 2).sum() / (6 * len(X))
    lt.backward()
    gt = Pt.grad.numpy()
    assert abs(lt.item() - loss) <= 1e-12 * abs(loss)
    assert np.abs(gt - g).max() <= 1e-10 * np.abs(gt).max()

    # Jacobian of raw outputs w.r.t. the six raw strain inputs
    _, _, J = mo.predict([P], nm, X, emulate=False, jacobian=True)
    Xt = torch.tensor(X.astype(np.float64), requires_grad=True)
    mean = torch.tensor(nm.in_mean.astype(np.float64))
    inv = torch.tensor(nm.in_inv.astype(np.float64))
    ostd = torch.tensor(nm.out_std.astype(np.float64))
    omean = torch.tensor(nm.out_mean.astype(np.float64))
    Pd = Pt.detach()
    for s in range(3):
        f = lambda x: net(Pd, ((x - mean) * inv)[None])[0] * ostd + omean
        Jt = torch.autograd.functional.jacobian(f, Xt[s].detach()).numpy()
        # the oracle normalises x in float32 as the kernel does: 1e-6 covers
        # that, a wrong index or transposition does not pass
        assert np.abs(Jt[:, :6] - J[s]).max() <= 1e-6 * np.abs(Jt).max()


def test_jacobian_matches_finite_differences():
    # identity normalisation and inputs on a 2^-20 grid, so x +- h is exact
    # in the float32 normalisation the oracle shares with the kernel
    P = [random_params(8), random_params(11)]
    nm = mo.Norm.identity()
    rng = np.random.default_rng(10)
    X = np.round(rng.normal(0, 1, (8, mo.IN)) * 2.0 ** 20) / 2.0 ** 20
    _, _, J = mo.predict(P, nm, X, emulate=False, jacobian=True)
    h = 2.0 ** -10
    for j in range(6):
        Xp, Xm = X.copy(), X.copy()
        Xp[:, j] += h
        Xm[:, j] -= h
        fd = (mo.predict(P, nm, Xp, emulate=False)[0] -
              mo.predict(P, nm, Xm, emulate=False)[0]) / (2 * h)
        assert np.abs(fd - J[:, :, j]).max() <= 1e-5 * np.abs(J).max()


def test_emulation_rounds_but_stays_close():
    P = random_params(12)
    nm = random_norm(13)
    rng = np.random.default_rng(14)
    X = rng.normal(0, 1, (128, mo.IN))
    Y = rng.normal(0, 1, (128, mo.OUT))
    y64 = mo.predict([P], nm, X, emulate=False)[0]
    ybf = mo.predict([P], nm, X, emulate=True)[0]
    d = np.abs(y64 - ybf).max() / np.abs(y64).max()
    assert 1e-5 < d < 5e-2
    l64, g64 = mo.loss_and_grad(P, nm, X, Y, emulate=False)
    lbf, gbf = mo.loss_and_grad(P, nm, X, Y, emulate=True)
    assert abs(l64 - lbf) < 5e-2 * l64
    assert np.linalg.norm(g64 - gbf) < 5e-2 * np.linalg.norm(g64)


def test_ensemble_mean_and_spread():
    Ps = [random_params(20 + k) for k in range(3)]
    nm = random_norm(30)
    X = np.random.default_rng(31).normal(0, 1, (16, mo.IN))
    ys = np.array([mo.predict([P], nm, X, emulate=False)[0] for P in Ps])
    Y, sd, _ = mo.predict(Ps, nm, X, emulate=False)
    assert np.allclose(Y, ys.mean(0), rtol=1e-12, atol=1e-12)
    assert np.allclose(sd, ys.std(0), rtol=1e-9, atol=1e-12)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(40)
    P = rng.normal(0, 1, 100)
    g = rng.normal(0, 1, 100)
    P1, m, v = mo.adam_step(P, np.zeros(100), np.zeros(100), g, 1, 1e-3, eps=0.0)
    assert np.allclose(P - P1, 1e-3 * np.sign(g), rtol=1e-12)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)
