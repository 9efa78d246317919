"""FE^2 message adapter, host side: the scale conversion of
MultiscaleCoupling.cc:7-22 on networks with known fiber lengths, and the
record checks of updateCoupling (MultiscaleRVEAnalysis.cc:144-180), which
reject unsupported RVE types and solvers before any device work."""
import math

import numpy as np
import pytest

import mumfim_amd as mf


def two_fiber_network():
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0]])
    fib = np.array([[0, 1], [1, 2]], np.int32)
    return mf.Network(X, fib, np.array([0, 2], np.int32),
                      [mf.Reaction(0, 0.1, 1.0, 1.0, 0, 0, 1.0)], np.zeros(2, np.int32))


def test_scale_conversion_known_lengths():
    net = two_fiber_network()            # total fiber length 2
    r = math.sqrt(0.25 / math.pi)        # fiber area 0.25
    assert mf.fo_scale_conversion(net, r, 0.5) == pytest.approx(1.0, rel=1e-15)
    assert mf.fo_scale_conversion(net, r, 2.0) == pytest.approx(4.0, rel=1e-15)
    # twice the fiber length (every fiber doubled) halves the conversion
    dbl = mf.Network(net.coords, np.concatenate([net.fibers, net.fibers]), net.boundary,
                     net.reactions, np.zeros(4, np.int32))
    assert mf.fo_scale_conversion(dbl, r, 0.5) == pytest.approx(0.5, rel=1e-15)
    with pytest.raises(mf.MfError):
        mf.fo_scale_conversion(net, 0.0, 0.5)


def _records(n, rve_type=mf.MF_FO_FIBER_ONLY, solver=mf.MF_FO_SOLVER_EXPLICIT):
    hdrs = np.zeros((n, 7), np.int32)
    hdrs[:, 0] = rve_type
    prms = np.zeros((n, 8))
    prms[:, 0] = 3.5e-8
    prms[:, 1] = 0.003
    prms[:, 7] = 1.0
    slvr = np.zeros((n, 8))
    slvr[:, 5] = 100.0
    slvr[:, 6] = 0.8
    sint = np.zeros((n, 10), np.int32)
    sint[:, 0] = solver
    return hdrs, prms, slvr, sint


@pytest.mark.parametrize("rve_type,solver,text", [
    (mf.MF_FO_ISOTROPIC_NEOHOOKEAN, mf.MF_FO_SOLVER_EXPLICIT, "Neohookean"),
    (7, mf.MF_FO_SOLVER_EXPLICIT, "not valid"),
    (mf.MF_FO_FIBER_ONLY, 0, "explicit"),
])
def test_adapter_rejects_unsupported_records(net27, rve_type, solver, text):
    with pytest.raises(mf.MfError, match=text):
        mf.MultiscaleRVEAdapter([net27], *_records(1, rve_type, solver))
