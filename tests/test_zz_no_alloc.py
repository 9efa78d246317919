"""The allocation-free run (mf_batch_run_device_noalloc) against the
allocating device run (mf_batch_run_device_async), through the C ABI: the
same stress bits, statuses and committed state, on two calls in a row, the
second on a smaller batch with a broadcast F and no status buffer.  The
first batch holds an RVE beyond the LDS kernels, so its global state lives
in the library's scratch array.  Sizes beyond the array and energy
diagnostics are refused before anything runs."""
import sys

import numpy as np
import pytest

import mumfim_amd as mf

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def strain_F(seed, n, amp):
    u = mf.uniform(seed, 77, 9 * n).reshape(n, 3, 3)
    U = (2 * u - 1) * amp
    return np.eye(3)[None] + 0.5 * (U + U.transpose(0, 2, 1))


def run(g, fn, F, update, with_status=True):
    """One run through entry point fn; returns stress, statuses (or None)."""
    F_dev = mf.DeviceArray.from_host(np.ascontiguousarray(F.reshape(-1, 9)))
    sig = mf.DeviceArray((g.n, 6))
    st = mf.DeviceArray((g.n,), np.int32) if with_status else None
    mf.check(fn(g.h, F_dev.data_ptr(), F.reshape(-1, 9).shape[0], sig.data_ptr(),
                int(update), st.data_ptr() if st is not None else None))
    g.synchronize()
    return sig.numpy(), (st.numpy() if st is not None else None)


@pytest.mark.parametrize("prec", [mf.MF_PREC_FP64, mf.MF_PREC_FP32C])
def test_noalloc_run_matches_device_run(prec):
    lib = mf.lib()
    kw = dict(mode=mf.MF_MODE_RAMP, n_steps=400, precision=prec,
              reaction_policy=mf.MF_REACT_PER_FIBER)
    # call 1: a 6000-node RVE (global-state kernel) beside two small ones,
    # one F per RVE, update run (commits the state)
    big = [mf.generate_network(seed=510 + i, n_nodes=n, nonlinear_fraction=0.5)
           for i, n in enumerate([6000, 400, 700])]
    F1 = strain_F(21, 3, 0.05)
    say(f"case: noalloc precision={prec}, call 1: 3 RVEs (6000/400/700 nodes), update")
    ga = mf.BatchedFiberRVEAnalysisExplicit(big, mf.default_params(**kw))
    gn = mf.BatchedFiberRVEAnalysisExplicit(big, mf.default_params(**kw))
    _, plan = gn.plan()
    assert plan[0, 2] & 8192, plan  # the big RVE runs on the global-state kernel
    sa, sta = run(ga, lib.mf_batch_run_device_async, F1, True)
    sn, stn = run(gn, lib.mf_batch_run_device_noalloc, F1, True)
    assert (sta == 0).all() and np.array_equal(stn, sta)
    assert np.array_equal(sn, sa)
    xa, ua, Va = ga.get_state()
    xn, un, Vn = gn.get_state()
    assert np.array_equal(xn, xa) and np.array_equal(un, ua) and np.array_equal(Vn, Va)
    # the committed state serves the tangent like a regular run's
    assert np.array_equal(gn.computeMaterialStiffness(), ga.computeMaterialStiffness())
    # call 2: a smaller batch, one broadcast F, statuses left in the scratch
    small = [mf.generate_network(seed=520 + i, n_nodes=300) for i in range(2)]
    F2 = strain_F(22, 1, 0.05)
    say(f"case: noalloc precision={prec}, call 2: 2 RVEs of 300 nodes, broadcast F, trial")
    ga2 = mf.BatchedFiberRVEAnalysisExplicit(small, mf.default_params(**kw))
    gn2 = mf.BatchedFiberRVEAnalysisExplicit(small, mf.default_params(**kw))
    sa2, sta2 = run(ga2, lib.mf_batch_run_device_async, F2, False)
    sn2, _ = run(gn2, lib.mf_batch_run_device_noalloc, F2, False, with_status=False)
    assert (sta2 == 0).all()
    assert np.array_equal(sn2, sa2)
    # and the first batch again (its trial from the committed state)
    sa3, _ = run(ga, lib.mf_batch_run_device_async, F1, False)
    sn3, _ = run(gn, lib.mf_batch_run_device_noalloc, F1, False)
    assert np.array_equal(sn3, sa3)
    for g in (ga, gn, ga2, gn2):
        g.close()


def test_noalloc_refuses_what_it_cannot_hold():
    lib = mf.lib()
    say("case: noalloc refusals")
    F = np.eye(3)[None]
    # energy diagnostics write a library buffer
    net = mf.generate_network(seed=530, n_nodes=200)
    g = mf.BatchedFiberRVEAnalysisExplicit(
        [net], mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=10, energy_diagnostics=1))
    F_dev = mf.DeviceArray.from_host(F.reshape(1, 9))
    sig = mf.DeviceArray((1, 6))
    rc = lib.mf_batch_run_device_noalloc(g.h, F_dev.data_ptr(), 1, sig.data_ptr(), 0, None)
    assert rc == -5  # MF_ERR_UNSUPPORTED
    g.close()
    # 64 RVEs of 12000 nodes need ~74 MB of global state: beyond the array
    gen = mf.default_gen_params(seed=540, n_nodes=12000)
    g = mf.BatchedFiberRVEAnalysisExplicit(
        params=mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=10), generated=(gen, 64))
    sig64 = mf.DeviceArray((64, 6))
    rc = lib.mf_batch_run_device_noalloc(g.h, F_dev.data_ptr(), 1, sig64.data_ptr(), 0, None)
    assert rc == -1  # MF_ERR_INVALID
    assert b"scratch" in lib.mf_last_error()
    g.close()
