"""BF16-MFMA surrogate kernels (through the C ABI) against the numpy oracle
(oracle/mlp_oracle.py): inference, the fused input Jacobian, the ensemble
statistics, the training gradients, the Adam update and training itself.

Tolerances: the oracle rounds to BF16 at the same points as the kernels and
accumulates in float64, so the kernels differ from it only by FP32
accumulation order and the rare BF16 rounding flip that order causes:
outputs and Jacobians within 2e-3 of their largest magnitude, gradients
within 2e-3 in relative L2 norm per parameter block.  Against the unrounded
float64 network (the mathematics) the bar is 5e-2.
"""
import math
import sys

import numpy as np
import pytest

import mumfim_amd as mf
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

BLOCKS = [("W1", mo.OFF_W1, mo.OFF_B1), ("b1", mo.OFF_B1, mo.OFF_W2),
          ("W2", mo.OFF_W2, mo.OFF_B2), ("b2", mo.OFF_B2, mo.OFF_W3),
          ("W3", mo.OFF_W3, mo.OFF_B3), ("b3", mo.OFF_B3, mo.OFF_W4),
          ("W4", mo.OFF_W4, mo.OFF_B4), ("b4", mo.OFF_B4, mo.NUM_PARAMS)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def setup_norm(m, seed, zero_std=True):
    rng = np.random.default_rng(seed)
    in_mean = rng.normal(0, 0.1, 10).astype(np.float32)
    in_std = rng.uniform(0.5, 2.0, 10).astype(np.float32)
    if zero_std:
        in_std[7] = 0.0  # constant descriptor: counts as 1
    out_mean = rng.normal(0, 0.1, 6).astype(np.float32)
    out_std = rng.uniform(0.5, 2.0, 6).astype(np.float32)
    m.set_normalization(in_mean, in_std, out_mean, out_std)
    return mo.Norm(in_mean, in_std, out_mean, out_std)


def rel_max(a, b):
    return float(np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-30))


def test_mlp_sizes_and_weights_roundtrip():
    say("case: mlp create/get/set weights")
    m = mf.SurrogateMLP(n_members=3, seed=1)
    assert mf.lib().mf_mlp_num_params() == mo.NUM_PARAMS
    assert mf.lib().mf_mlp_num_members(m.h) == 3
    w = [m.get_weights(k) for k in range(3)]
    assert not np.array_equal(w[0], w[1])  # members initialised independently
    bound = 1 / np.sqrt(10)
    assert np.abs(w[0][mo.OFF_W1:mo.OFF_B1]).max() <= bound
    assert np.abs(w[0][mo.OFF_W2:mo.OFF_B2]).max() <= 1 / 16
    w2 = np.random.default_rng(0).normal(0, 0.05, mo.NUM_PARAMS).astype(np.float32)
    m.set_weights(1, w2)
    assert np.array_equal(m.get_weights(1), w2)
    assert np.array_equal(m.get_weights(0), w[0])
    with pytest.raises(mf.MfError):
        m.get_weights(3)


@pytest.mark.parametrize("n", [1, 300, 1024])
def test_mlp_inference_and_jacobian_match_oracle(n):
    say(f"case: mlp inference n={n}")
    m = mf.SurrogateMLP(n_members=2, seed=3)
    nm = setup_norm(m, 4)
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (n, 10)).astype(np.float32) * 2
    Ps = [m.get_weights(k) for k in range(2)]
    for member in (0, 1, -1):
        Pl = Ps if member < 0 else [Ps[member]]
        Yj, Sj, J = m.predict(X, jacobian=True, member=member)
        Y, S, _ = m.predict(X, jacobian=False, member=member)
        # the Jacobian kernel computes plane 0 with the same products as the
        # plain kernel (the compiler may contract the epilogues differently)
        assert rel_max(Y, Yj) <= 1e-6
        Yo, So, Jo = mo.predict(Pl, nm, X, emulate=True, jacobian=True)
        Y64, _, J64 = mo.predict(Pl, nm, X, emulate=False, jacobian=True)
        ey, ej = rel_max(Y, Yo), rel_max(J, Jo)
        say(f"  member {member}: y {ey:.2e} J {ej:.2e} | vs float64 y {rel_max(Y, Y64):.2e} "
            f"J {rel_max(J, J64):.2e}")
        assert ey <= 2e-3 and ej <= 2e-3
        assert rel_max(Y, Y64) <= 5e-2 and rel_max(J, J64) <= 5e-2
        if member < 0:
            assert np.abs(S - So).max() <= 2e-3 * np.abs(Yo).max()
            assert np.abs(S - Sj).max() <= 1e-6 * np.abs(Y).max()
        else:
            assert np.all(S == 0)


def test_mlp_gradient_matches_oracle():
    say("case: mlp gradient")
    m = mf.SurrogateMLP(n_members=2, seed=7)
    nm = setup_norm(m, 8)
    rng = np.random.default_rng(9)
    n = 512
    X = rng.normal(0, 1, (n, 10)).astype(np.float32)
    Y = rng.normal(0, 1, (n, 6)).astype(np.float32)
    Xd, Yd = mf.DeviceArray.from_host(X), mf.DeviceArray.from_host(Y)
    for member in (0, 1):
        P = m.get_weights(member)
        loss, g = m.gradient(member, Xd, Yd, n)
        lo, go = mo.loss_and_grad(P, nm, X, Y, emulate=True)
        l64, g64 = mo.loss_and_grad(P, nm, X, Y, emulate=False)
        assert abs(loss - lo) <= 1e-4 * lo, (loss, lo)
        for name, a, b in BLOCKS:
            e = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
            e64 = np.linalg.norm(g[a:b] - g64[a:b]) / np.linalg.norm(g64[a:b])
            say(f"  member {member} {name}: {e:.2e} (vs float64 {e64:.2e})")
            assert e <= 2e-3, name
            assert e64 <= 5e-2, name


def test_adam_steps_match_oracle():
    say("case: adam steps")
    m = mf.SurrogateMLP(n_members=1, seed=11)
    setup_norm(m, 12)
    rng = np.random.default_rng(13)
    n = 256
    X = rng.normal(0, 1, (n, 10)).astype(np.float32)
    Y = rng.normal(0, 1, (n, 6)).astype(np.float32)
    Xd, Yd = mf.DeviceArray.from_host(X), mf.DeviceArray.from_host(Y)
    P = m.get_weights(0).astype(np.float64)
    mom, vel = np.zeros_like(P), np.zeros_like(P)
    for t, lr in ((1, 1e-3), (2, 5e-4)):
        _, g = m.gradient(0, Xd, Yd, n)
        m.train_step(Xd, Yd, n, n, lr)
        P, mom, vel = mo.adam_step(P, mom, vel, g.astype(np.float64), t, lr)
        got = m.get_weights(0)
        err = np.abs(got - P).max()
        say(f"  step {t}: max |dP| {err:.2e}")
        assert err <= 1e-7 + 1e-3 * lr


def test_index_gather_equals_gathered_batch():
    say("case: train step with row indices")
    rng = np.random.default_rng(14)
    N, B = 1024, 256
    X = rng.normal(0, 1, (N, 10)).astype(np.float32)
    Y = rng.normal(0, 1, (N, 6)).astype(np.float32)
    idx = np.stack([rng.permutation(N)[:B] for _ in range(2)]).astype(np.int32)
    a = mf.SurrogateMLP(n_members=2, seed=15)
    b = mf.SurrogateMLP(n_members=1, seed=15)  # member 0 of a == member 0 of b
    for mm in (a, b):
        setup_norm(mm, 16)
    la = a.train_step(mf.DeviceArray.from_host(X), mf.DeviceArray.from_host(Y), N, B, 1e-3,
                      idx_dev=mf.DeviceArray.from_host(idx))
    b.train_step(mf.DeviceArray.from_host(X[idx[0]]), mf.DeviceArray.from_host(Y[idx[0]]), B, B,
                 1e-3)
    assert np.isfinite(la)
    assert np.array_equal(a.get_weights(0), b.get_weights(0))
    assert not np.array_equal(a.get_weights(1), b.get_weights(0))


def _smooth_target(X, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(0, 0.5, (10, 16))
    B = rng.normal(0, 1.0, (16, 6))
    return (np.tanh(X @ A) @ B + 0.3 * X[:, :6] ** 2).astype(np.float32)


def test_training_fits_a_smooth_map_and_is_deterministic():
    say("case: fit smooth map")
    rng = np.random.default_rng(17)
    n = 16384
    X = rng.normal(0, 1, (n, 10)).astype(np.float32)
    Y = _smooth_target(X, 18)
    Xd, Yd = mf.DeviceArray.from_host(X), mf.DeviceArray.from_host(Y)
    ws = []
    for rep in range(2):
        m = mf.SurrogateMLP(n_members=2, seed=19)
        m.set_normalization(X.mean(0), X.std(0), Y.mean(0), Y.std(0))
        hist = m.fit(Xd, Yd, n, epochs=12, batch=256, lr=2e-3, lr_final=1e-4, seed=20)
        say(f"  rep {rep}: epoch losses {hist.mean(1).round(5).tolist()}")
        assert hist[-1].mean() < 0.05 * hist[0].mean()
        ws.append(m.get_weights(1))
    assert np.array_equal(ws[0], ws[1])  # fixed-order reductions: reproducible
    Xt = rng.normal(0, 1, (4096, 10)).astype(np.float32)
    Yt = _smooth_target(Xt, 18)
    Yp, S, _ = m.predict(Xt)
    rms = np.sqrt(((Yp - Yt) ** 2).mean()) / Yt.std()
    say(f"  held-out relative rms {rms:.3e}, mean spread {S.mean():.3e}")
    assert rms < 0.1


def _splitmix_perm(seed, stream, n):
    """The fit's per-epoch, per-member shuffle: Fisher-Yates driven by
    SplitMix64(seed, stream) (the library's generator)."""
    M = (1 << 64) - 1

    def mix(z):
        z = (z + 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    s = mix(seed ^ mix(stream))
    p = list(range(n))
    for i in range(n - 1, 0, -1):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        j = z % (i + 1)
        p[i], p[j] = p[j], p[i]
    return np.array(p, np.int32)


def test_fit_equals_its_steps():
    """fit over three epochs gives bitwise the weights of the same Adam steps
    launched one by one through train_step, with the fit's per-epoch
    SplitMix64 shuffles and its cosine learning-rate schedule restated here:
    pins the shuffle, the schedule and the bias-correction step count."""
    say("case: fit vs its steps")
    rng = np.random.default_rng(27)
    n, B, epochs, K, seed = 2048, 256, 3, 2, 31
    lr0, lr1 = 2e-3, 1e-4
    X = rng.normal(0, 1, (n, 10)).astype(np.float32)
    Y = _smooth_target(X, 28)
    Xd, Yd = mf.DeviceArray.from_host(X), mf.DeviceArray.from_host(Y)
    a = mf.SurrogateMLP(n_members=K, seed=29)
    b = mf.SurrogateMLP(n_members=K, seed=29)
    for m in (a, b):
        m.set_normalization(X.mean(0), X.std(0), Y.mean(0), Y.std(0))
    tp = mf.default_train_params()
    assert (tp.beta1, tp.beta2) == (np.float32(0.9), np.float32(0.999))
    hist = a.fit(Xd, Yd, n, epochs=epochs, batch=B, lr=lr0, lr_final=lr1, seed=seed)
    steps = n // B
    total = steps * epochs
    for ep in range(epochs):
        perms = [_splitmix_perm(seed, ep * K + k, n) for k in range(K)]
        for st in range(steps):
            it = ep * steps + st
            frac = it / (total - 1)
            # the schedule as the library evaluates it: lr - lr_final in FP32,
            # the rest in double, rounded to FP32
            l1 = float(np.float32(lr1))
            dl = float(np.float32(lr0) - np.float32(lr1))
            lr = float(np.float32(l1 + 0.5 * dl * (1.0 + math.cos(math.pi * frac))))
            idx = np.concatenate([p[st * B:(st + 1) * B] for p in perms]).astype(np.int32)
            b.train_step(Xd, Yd, n, B, lr, idx_dev=mf.DeviceArray.from_host(idx),
                         return_loss=False)
    for k in range(K):
        assert np.array_equal(a.get_weights(k), b.get_weights(k)), k
    say(f"  epoch losses {hist.mean(1).round(5).tolist()}")
    assert hist[-1].mean() < hist[0].mean()
