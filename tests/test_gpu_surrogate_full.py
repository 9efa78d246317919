"""The learned surrogate at its full workload: 1,048,576 training solves
(65536 generated networks of 500 nodes x 16 strains, F = I + sym(U),
U ~ U(-0.1, 0.1), 5000 ramp steps, FP32c explicit solver), a 4-member
ensemble trained on the GPU, inference on 2^20 points with and without the
tangent, kernels checked against the numpy oracle on a 4096-point sample,
and the explicit fallback of the surrogate path checked bitwise on a
held-out batch of 16384 networks."""
import sys
import time

import numpy as np
import pytest

import mumfim_amd as mf
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

N_NET, PER, STEPS = 65536, 16, 5000


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def strain_F(seed, n, amp):
    u = mf.uniform(seed, 77, 9 * n).reshape(n, 3, 3)
    U = (2 * u - 1) * amp
    return np.eye(3)[None] + 0.5 * (U + U.transpose(0, 2, 1))


def rel_max(a, b):
    return float(np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-30))


def params():
    return mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=STEPS, precision=mf.MF_PREC_FP32C,
                             reaction_policy=mf.MF_REACT_PER_FIBER)


def test_surrogate_full_size():
    say(f"case: surrogate at full size, {N_NET} networks x {PER} strains")
    t0 = time.perf_counter()
    g = mf.BatchedFiberRVEAnalysisExplicit(params=params(),
                                           generated=(mf.default_gen_params(seed=1), N_NET))
    say(f"  networks built in {time.perf_counter() - t0:.1f}s")
    n = N_NET * PER
    X = np.empty((n, mf.MF_MLP_IN), np.float32)
    Y = np.empty((n, 6), np.float32)
    F_dev = mf.DeviceArray((N_NET, 9))
    sig_dev = mf.DeviceArray((N_NET, 6))
    st_dev = mf.DeviceArray((N_NET,), np.int32)
    for k in range(PER):
        F = strain_F(1000 + k, N_NET, 0.1)
        F_dev.copy_from(F.reshape(N_NET, 9))
        g.run_device(F_dev, sig_dev, status_dev=st_dev)
        g.synchronize()
        assert (st_dev.numpy() == mf.MF_STATUS_OK).all()
        Y[k * N_NET:(k + 1) * N_NET] = sig_dev.numpy()
        X[k * N_NET:(k + 1) * N_NET] = mf.surrogate_inputs(g, F)
        if k % 4 == 3:
            say(f"  training data {N_NET * (k + 1)} solves ({time.perf_counter() - t0:.1f}s)")
    assert np.isfinite(Y).all()
    g.close()
    del F_dev, sig_dev, st_dev

    Xd, Yd = mf.DeviceArray.from_host(X), mf.DeviceArray.from_host(Y)
    m = mf.SurrogateMLP(n_members=4, seed=7)
    m.set_normalization(X.mean(0), X.std(0), Y.mean(0), Y.std(0))
    hist = m.fit(Xd, Yd, n, epochs=30, batch=4096, lr=2e-3, lr_final=1e-5, seed=8)
    say(f"  trained 30 epochs: loss {hist[0].mean():.4f} -> {hist[-1].mean():.4f}")
    assert hist[-1].mean() < 0.9 * hist[0].mean()

    # inference on 2^20 points, without and with the fused tangent
    npts = 1 << 20
    Yp = mf.DeviceArray((npts, 6), np.float32)
    Sp = mf.DeviceArray((npts, 6), np.float32)
    Yj = mf.DeviceArray((npts, 6), np.float32)
    Sj = mf.DeviceArray((npts, 6), np.float32)
    Jp = mf.DeviceArray((npts, 6, 6), np.float32)
    m.predict_device(Xd, npts, Yp, None, Sp)
    ms_sig = m.last_predict_ms()
    m.predict_device(Xd, npts, Yj, Jp, Sj)
    ms_jac = m.last_predict_ms()
    y, s, yj, sj, J = Yp.numpy(), Sp.numpy(), Yj.numpy(), Sj.numpy(), Jp.numpy()
    say(f"  2^20 points: {ms_sig:.2f} ms stress, {ms_jac:.2f} ms stress + tangent")
    assert np.isfinite(y).all() and np.isfinite(J).all()
    assert rel_max(yj, y) <= 1e-6 and rel_max(sj, s) <= 1e-6
    idx = np.sort(np.random.default_rng(11).choice(npts, 4096, replace=False))
    Ps = [m.get_weights(k) for k in range(4)]
    norm = mo.Norm(*m.normalization())
    Yo, So, Jo = mo.predict(Ps, norm, X[idx], emulate=True, jacobian=True)
    ey, ej = rel_max(y[idx], Yo), rel_max(J[idx], Jo)
    es = float(np.abs(s[idx] - So).max() / np.abs(Yo).max())
    say(f"  4096 sampled points vs oracle: y {ey:.2e}  J {ej:.2e}  spread {es:.2e}")
    assert ey <= 2e-3 and ej <= 2e-3 and es <= 2e-3
    del Xd, Yd, Yp, Sp, Yj, Sj, Jp

    # surrogate path on held-out networks: rejected RVEs get the explicit
    # solver's stress bitwise
    nt = 16384
    gt = mf.BatchedFiberRVEAnalysisExplicit(
        params=params(), generated=(mf.default_gen_params(seed=20_000_001), nt))
    Ft = strain_F(3001, nt, 0.1)
    s_ref, st_ref = gt.run(Ft, update_coords=False, return_status=True)
    assert (st_ref == mf.MF_STATUS_OK).all()
    Xin = mf.surrogate_inputs(gt, Ft)
    yt, spt, _ = m.predict(Xin)
    q = spt.max(1) / np.abs(yt).max(1)
    tol = float(np.median(q))
    m.set_estimator(X.min(0), X.max(0), margin=0.02, kappa=1.0, sigma_floor=0.0, tol=tol)
    sig, _, st, stats = mf.surrogate_run(gt, m, Ft)
    fb = st == mf.MF_STATUS_FALLBACK
    say(f"  held-out {nt}: {stats['n_fallback']} fallbacks "
        f"({stats['n_out_of_range']} out of range), path {stats['total_ms']:.1f} ms")
    assert 0 < stats["n_fallback"] < nt and fb.sum() == stats["n_fallback"]
    assert np.all((st == mf.MF_STATUS_OK) | fb)
    assert np.array_equal(sig[fb], s_ref[fb])
    assert np.array_equal(sig[~fb], yt[~fb].astype(np.float64))
