"""FE^2 message adapter on the GPU: one macro iteration is run(F) with update
plus the tangent, packed as micro_fo_result; a macro step ends with the
orientation tensors as micro_fo_step_result; the per-RVE damping and scale
factor the records carry reach the solve.  Checked bitwise against plain
batches configured by hand, and the stress against the CPU oracle."""
import copy
import sys

import numpy as np
import pytest

import mumfim_amd as mf
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def records(n, eta):
    hdrs = np.zeros((n, 7), np.int32)
    hdrs[:, 0] = mf.MF_FO_FIBER_ONLY
    prms = np.zeros((n, 8))
    prms[:, 0] = 3.5e-8          # FIBER_RADIUS
    prms[:, 1] = 0.003           # VOLUME_FRACTION
    prms[:, 5:8] = [0.0, 0.6, 0.8]  # ORIENTATION_AXIS
    slvr = np.zeros((n, 8))
    slvr[:, 5] = eta             # VISCOUS_DAMPING_FACTOR
    slvr[:, 6] = 0.8
    sint = np.zeros((n, 10), np.int32)
    sint[:, 0] = mf.MF_FO_SOLVER_EXPLICIT
    return hdrs, prms, slvr, sint


def test_adapter_matches_hand_configured_batches(net27):
    say("case: FE2 adapter, 2 RVEs of Network_2_7, damping 100 and 60")
    eta = np.array([100.0, 60.0])
    ad = mf.MultiscaleRVEAdapter([net27, net27], *records(2, eta))
    scale = mf.fo_scale_conversion(net27, 3.5e-8, 0.003)
    F1 = np.array([[1.01, 0.002, 0.0], [0.0, 0.995, 0.001], [0.0, 0.0, 1.0]])
    F2 = np.array([[1.005, 0.0, 0.0], [0.001, 1.0, 0.0], [0.0, 0.0, 0.998]])
    data1 = np.stack([F1.ravel(), F2.ravel()])
    data2 = np.stack([F2.ravel(), F1.ravel()])
    res1, st1 = ad.run(data1, return_status=True)
    res2 = ad.run(data2)
    step = ad.step_results()
    assert np.all(st1 == mf.MF_STATUS_OK)
    net_s = copy.copy(net27)
    net_s.scale_factor = scale
    for r in range(2):
        g = mf.BatchedFiberRVEAnalysisExplicit([net_s], mf.default_params(visc_damp_coeff=eta[r]))
        for data, res in ((data1, res1), (data2, res2)):
            s = g.run(data[r].reshape(1, 3, 3))
            Cm = g.computeMaterialStiffness()
            assert np.array_equal(res[r, :6], s[0])
            assert np.all(res[r, 6:9] == 0)
            assert np.array_equal(res[r, 9:], Cm[0].ravel())
        assert np.array_equal(step[r, :9], g.compute3DOrientationTensor()[0].ravel())
        assert np.array_equal(step[r, 9:], g.compute2DOrientationTensor([0.0, 0.6, 0.8])[0].ravel())
    # stress against the CPU oracle (scale factor included)
    o = Oracle([net_s], mf.default_params(visc_damp_coeff=eta[0]))
    so = o.run(data1[0].reshape(1, 3, 3))
    err = np.abs(res1[0, :6] - so[0]).max() / np.abs(so[0]).max()
    say(f"  sigma vs oracle {err:.2e}, scale {scale:.6e}")
    assert err <= 1e-10


def test_per_rve_damping_equals_separate_batches():
    say("case: per-RVE damping, 4 generated RVEs, FP64 and FP32c ramp")
    g = mf.default_gen_params(seed=7, n_nodes=300)
    eta = np.array([100.0, 40.0, 250.0, 100.0])
    F = np.tile(np.diag([1.03, 0.99, 1.0]), (4, 1, 1))
    for prec in (mf.MF_PREC_FP64, mf.MF_PREC_FP32C):
        p = mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=1500, precision=prec)
        b = mf.BatchedFiberRVEAnalysisExplicit(params=p, generated=(g, 4))
        b.set_damping(eta)
        s = b.run(F, update_coords=False)
        for r in range(4):
            pr = mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=1500, precision=prec,
                                   visc_damp_coeff=eta[r])
            gr = mf.default_gen_params(seed=7 + r, n_nodes=300)
            one = mf.BatchedFiberRVEAnalysisExplicit(params=pr, generated=(gr, 1))
            assert np.array_equal(s[r], one.run(F[r:r + 1], update_coords=False)[0])
        b.set_damping(None)
        s0 = b.run(F, update_coords=False)
        assert np.array_equal(s0[0], s[0]) and not np.array_equal(s0[1], s[1])
    with pytest.raises(mf.MfError):
        b.set_damping([1.0, -1.0, 0.0, 0.0])
