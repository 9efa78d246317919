"""The surrogate path with explicit fallback (mf_surrogate_run) on batches of
generated networks: an ensemble trained on explicit solves of the same
generator predicts held-out RVEs far better than an untrained one and much
faster than the explicit solver; rejected RVEs get exactly the explicit
solver's stress and central tangent; out-of-range inputs are rejected; the
surrogate tangent is the ensemble's input Jacobian mapped to the
perturbations F_i(h) F."""
import sys

import numpy as np
import pytest

import mumfim_amd as mf

pytestmark = pytest.mark.gpu

STEPS = 1000


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def strain_F(seed, n, amp):
    u = mf.uniform(seed, 77, 9 * n).reshape(n, 3, 3)
    U = (2 * u - 1) * amp
    return np.eye(3)[None] + 0.5 * (U + U.transpose(0, 2, 1))


def batch(seed0, n, prec=None):
    g = mf.default_gen_params(seed=seed0, n_nodes=500)
    p = mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=STEPS,
                          precision=mf.MF_PREC_FP32C if prec is None else prec,
                          reaction_policy=mf.MF_REACT_PER_FIBER,
                          tangent_scheme=mf.MF_TANGENT_CENTRAL, fd_step=1e-5)
    return mf.BatchedFiberRVEAnalysisExplicit(params=p, generated=(g, n))


def rel_err(a, b):
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    return np.abs(a - b).max(1) / np.maximum(np.abs(b).max(1), 1e-300)


@pytest.fixture(scope="module")
def trained():
    say("case: surrogate training data (1024 networks x 4 strains)")
    a = batch(1, 1024)
    Xs, Ys = [], []
    for k in range(4):
        F = strain_F(100 + k, a.n, 0.1)
        s = a.run(F, update_coords=False)
        Xs.append(mf.surrogate_inputs(a, F))
        Ys.append(s)
    X = np.concatenate(Xs).astype(np.float32)
    Y = np.concatenate(Ys).astype(np.float32)
    m = mf.SurrogateMLP(n_members=4, seed=5)
    m.set_normalization(X.mean(0), X.std(0), Y.mean(0), Y.std(0))
    hist = m.fit(mf.DeviceArray.from_host(X), mf.DeviceArray.from_host(Y), len(X),
                 epochs=40, batch=512, lr=2e-3, lr_final=2e-5, seed=6)
    say(f"  loss first/last epoch {hist[0].mean():.4f} / {hist[-1].mean():.4f}")
    m.set_estimator(X.min(0), X.max(0), margin=0.05, kappa=1.0, sigma_floor=0.0, tol=1e30)
    return m, X, Y


def test_trained_surrogate_beats_untrained_and_is_fast(trained):
    m = trained[0]
    a = batch(50001, 512)
    F = strain_F(900, a.n, 0.1)
    s_ref = a.run(F, update_coords=False)
    ms_explicit, _ = a.last_solve_ms()
    sig, _, st, stats = mf.surrogate_run(a, m, F)
    assert stats["n_fallback"] == 0 and np.all(st == mf.MF_STATUS_OK)
    untrained = mf.SurrogateMLP(n_members=4, seed=99)
    nrm = m.normalization()
    untrained.set_normalization(*nrm)
    su, _, _, _ = mf.surrogate_run(a, untrained, F)
    e, eu = rel_err(sig, s_ref), rel_err(su, s_ref)
    mean_sigma = np.tile(s_ref.mean(0), (a.n, 1))
    em = rel_err(mean_sigma, s_ref)
    # least-squares linear elasticity on the same training data
    _, Xtr, Ytr = trained
    A = np.c_[Xtr[:, :6], np.ones(len(Xtr))].astype(np.float64)
    coef = np.linalg.lstsq(A, Ytr.astype(np.float64), rcond=None)[0]
    Xin = mf.surrogate_inputs(a, F)
    el = rel_err(np.c_[Xin[:, :6], np.ones(a.n)] @ coef, s_ref)
    say(f"  rel err median/p90: trained {np.median(e):.3f}/{np.percentile(e, 90):.3f} "
        f"untrained {np.median(eu):.3f} batch-mean stress {np.median(em):.3f} "
        f"linear fit {np.median(el):.3f}")
    say(f"  surrogate {stats['mlp_ms']:.3f} ms vs explicit {ms_explicit:.3f} ms")
    assert np.median(e) < 0.6 * np.median(eu)
    assert np.median(e) < 0.6 * np.median(em)
    # 512 RVEs leave the surrogate's fixed costs (F upload, four small
    # launches) dominant; the bench measures it on a million points
    assert stats["mlp_ms"] * 4 < ms_explicit


def test_rejected_rves_get_the_explicit_solution(trained):
    m, X = trained[0], trained[1]
    a = batch(60001, 128)
    F = strain_F(901, a.n, 0.1)
    F[::5] = strain_F(902, len(F[::5]), 0.4)  # outside the training strains
    s_ref = a.run(F, update_coords=False)
    est = m.estimator()
    m.set_estimator(X.min(0), X.max(0), margin=0.05, kappa=1.0, sigma_floor=0.0, tol=1e30)
    sig, _, st, stats = mf.surrogate_run(a, m, F)
    oor = st == mf.MF_STATUS_FALLBACK
    say(f"  out of range: {stats['n_out_of_range']} of {a.n}")
    assert stats["n_out_of_range"] == stats["n_fallback"] == oor.sum() >= len(F[::5]) - 2
    assert np.array_equal(sig[oor], s_ref[oor])
    # spread threshold 0: every RVE falls back and matches the solver bitwise
    m.set_estimator(X.min(0), X.max(0), margin=10.0, kappa=1.0, sigma_floor=0.0, tol=0.0)
    sig0, _, st0, stats0 = mf.surrogate_run(a, m, F)
    assert stats0["n_fallback"] == a.n and np.all(st0 == mf.MF_STATUS_FALLBACK)
    assert np.array_equal(sig0, s_ref)
    # a threshold between: accepted RVEs carry the ensemble mean
    Xin = mf.surrogate_inputs(a, F)
    Yp, Sp, _ = m.predict(Xin)
    q = np.max(Sp, 1) / np.abs(Yp).max(1)
    tol = float(np.median(q))
    m.set_estimator(X.min(0), X.max(0), margin=10.0, kappa=1.0, sigma_floor=0.0, tol=tol)
    sig1, _, st1, stats1 = mf.surrogate_run(a, m, F)
    acc = st1 == mf.MF_STATUS_OK
    say(f"  tol {tol:.3e}: {stats1['n_fallback']} fallbacks")
    assert 0 < stats1["n_fallback"] < a.n
    assert np.array_equal(sig1[~acc], s_ref[~acc])
    assert np.allclose(sig1[acc], Yp[acc].astype(np.float64), rtol=0, atol=0)
    m.set_estimator(est.in_lo, est.in_hi, est.margin, est.kappa, est.sigma_floor, est.tol)


def _perturbation(i, h):
    D1 = np.sqrt(1 / (1 - 2 * h))
    l2 = (1 + h) / (1 - h * h)
    l3 = (1 - h) / (1 - h * h)
    p = 0.5 * (np.sqrt(l2) + np.sqrt(l3))
    q = 0.5 * (np.sqrt(l2) - np.sqrt(l3))
    Fp = np.eye(3)
    if i < 3:
        Fp[i, i] = D1
    else:
        a, b = {3: (1, 2), 4: (0, 2), 5: (0, 1)}[i]
        Fp[:, :] = np.eye(3) * p
        Fp[3 - a - b, 3 - a - b] = 1
        Fp[a, b] = Fp[b, a] = q
    return Fp


def test_surrogate_tangent(trained):
    m, X = trained[0], trained[1]
    a = batch(70001, 64, prec=mf.MF_PREC_FP64)
    F = strain_F(903, a.n, 0.05)
    # accepted: C = J e(sym(E_i F)) + W S - S W, recomputed here in numpy
    m.set_estimator(X.min(0), X.max(0), margin=10.0, kappa=1.0, sigma_floor=0.0, tol=1e30)
    sig, C, st, _ = mf.surrogate_run(a, m, F, tangent=True)
    Xin = mf.surrogate_inputs(a, F)
    Yp, _, J = m.predict(Xin, jacobian=True)
    vi, vj = [0, 1, 2, 1, 0, 0], [0, 1, 2, 2, 2, 1]
    Cn = np.zeros((a.n, 6, 6))
    for r in range(a.n):
        s = Yp[r].astype(np.float64)
        S = np.array([[s[0], s[5], s[4]], [s[5], s[1], s[3]], [s[4], s[3], s[2]]])
        for i in range(6):
            E = np.zeros((3, 3))
            if i < 3:
                E[i, i] = 1
            else:
                E[vi[i], vj[i]] = E[vj[i], vi[i]] = 0.5
            G = E @ F[r]
            de = np.array([G[0, 0], G[1, 1], G[2, 2], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0],
                           G[0, 1] + G[1, 0]])
            W = 0.5 * (G - G.T)
            R = W @ S - S @ W
            Cn[r, :, i] = J[r].astype(np.float64) @ de + np.array([R[vi[k], vj[k]] for k in range(6)])
    assert rel_err(C, Cn).max() < 1e-9
    # rejected: the explicit central tangent from the reference state
    m.set_estimator(X.min(0), X.max(0), margin=10.0, kappa=1.0, sigma_floor=0.0, tol=0.0)
    sig0, C0, st0, _ = mf.surrogate_run(a, m, F, tangent=True)
    h = 1e-5
    Cfd = np.zeros((a.n, 6, 6))
    for i in range(6):
        sp = a.run(np.einsum("ab,rbc->rac", _perturbation(i, h), F), update_coords=False)
        sm = a.run(np.einsum("ab,rbc->rac", _perturbation(i, -h), F), update_coords=False)
        Cfd[:, :, i] = (sp - sm) / (2 * h)
    e_fb = rel_err(C0, Cfd).max()
    e_sur = np.median(rel_err(C, Cfd))
    say(f"  fallback tangent vs central FD {e_fb:.2e}; surrogate tangent median err {e_sur:.3f}")
    assert e_fb < 1e-9
    assert np.all(st0 == mf.MF_STATUS_FALLBACK)


def test_surrogate_needs_reference_state(trained):
    m = trained[0]
    a = batch(80001, 8)
    F = strain_F(904, a.n, 0.05)
    a.run(F, update_coords=True)
    with pytest.raises(mf.MfError):
        mf.surrogate_run(a, m, F)
