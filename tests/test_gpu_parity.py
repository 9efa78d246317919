"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle on the
same seeded networks and deformation gradients.

Tolerances: FP64 path <= 1e-10 relative stress error (relative to the largest
stress component of each RVE); FP32 + compensated path <= 1e-4 relative.
"""
import dataclasses
import math
import os
import sys

import numpy as np
import pytest

import mumfim_amd as mf
from conftest import GOLDEN, voigt
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32C_TOL = 1e-4


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    """max over RVEs of max|a-b| / max|b| (per RVE)."""
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    den = np.maximum(np.abs(b).max(1), 1e-300)
    return float((np.abs(a - b).max(1) / den).max())


def strain_F(seed, n, amp):
    """F = I + sym(U), U_ij ~ U(-amp, amp) from the SplitMix64 stream."""
    u = mf.uniform(seed, 77, 9 * n).reshape(n, 3, 3)
    U = (2 * u - 1) * amp
    return np.eye(3)[None] + 0.5 * (U + U.transpose(0, 2, 1))


def ramp_params(**kw):
    kw.setdefault("mode", mf.MF_MODE_RAMP)
    kw.setdefault("n_steps", 5000)
    return mf.default_params(**kw)


# ---------------------------------------------------------------- reference mode
def test_network27_relax_fp64_matches_oracle(golden, net27):
    say("case: Network_2_7 relax FP64")
    F = np.array(golden["network_2_7"]["F"])
    o = Oracle([net27])
    s_o = o.run(F)
    g = mf.BatchedFiberRVEAnalysisExplicit([net27])
    s_g, st = g.run(F, return_status=True)
    assert st[0] == mf.MF_STATUS_OK
    err = rel_err(s_g, s_o)
    say(f"  sigma rel err {err:.3e}")
    assert err <= FP64_TOL
    xo, uo, Vo = o.get_state()
    xg, ug, Vg = g.get_state()
    say(f"  state max|dx| {np.abs(xg - xo).max():.3e} bitwise={np.array_equal(xg, xo)}")
    assert np.abs(xg - xo).max() <= 1e-14 and np.abs(ug - uo).max() <= 1e-14
    assert Vg[0] == Vo[0]
    # golden: non-batched stress / det F
    J = np.linalg.det(F)
    sg = np.array(voigt(golden["network_2_7"]["explicit_cpu"][0]["stress33"])) / J
    assert np.abs(s_g[0] - sg).max() <= 2e-6 * np.abs(sg).max()


def test_network27_tangent_fp64_matches_oracle(golden, net27):
    say("case: Network_2_7 tangent FP64")
    F = np.array(golden["network_2_7"]["F"])
    o = Oracle([net27])
    o.run(F)
    C_o = o.computeMaterialStiffness()
    g = mf.BatchedFiberRVEAnalysisExplicit([net27])
    g.run(F)
    C_g, st = g.computeMaterialStiffness(return_status=True)
    assert (st == 0).all()
    err = rel_err(C_g, C_o)
    say(f"  C rel err {err:.3e}")
    assert err <= 1e-8
    J = np.linalg.det(F)
    Cg = np.array(golden["network_2_7"]["explicit_cpu"][0]["C"]) / J
    assert np.abs(C_g[0] - Cg).max() / np.abs(Cg).max() < 5e-3
    # the trial runs left the same state behind (restored u, trial x)
    xo, uo, _ = o.get_state()
    xg, ug, _ = g.get_state()
    assert np.abs(xg - xo).max() <= 1e-14 and np.array_equal(ug, uo)
    # central scheme: 12 chained trials, the -h trial of a column starting
    # where its +h trial ended (the oracle's scheme, pinned by the chain
    # closed form)
    pc = mf.default_params(tangent_scheme=mf.MF_TANGENT_CENTRAL, fd_step=1e-5)
    oc = Oracle([net27], pc)
    oc.run(F)
    Cc_o = oc.computeMaterialStiffness()
    gc = mf.BatchedFiberRVEAnalysisExplicit([net27], pc)
    gc.run(F)
    Cc_g, stc = gc.computeMaterialStiffness(return_status=True)
    assert (stc == 0).all()
    errc = rel_err(Cc_g, Cc_o)
    say(f"  central C rel err {errc:.3e}")
    assert errc <= 1e-8
    # forward and central agree to the scheme's truncation and trial order
    assert rel_err(Cc_g, C_g) < 5e-3


def test_relax_toe_law_fp64_fiber_once_matches_oracle():
    """Reference (relax) mode on toe-law RVEs through the FP64 kernel that
    evaluates each fiber once per step: stress within the FP64 bar of the
    oracle and the committed state equal to rounding, like the node-centric
    FP64 kernel (whose results it reproduces bitwise)."""
    say("case: relax, toe law, FP64 fiber-once kernel")
    nets = [mf.generate_network(seed=23 + i, n_nodes=n, E=1e8, radius=5e-5,
                                nonlinear_fraction=0.5) for i, n in enumerate([200, 300])]
    F = strain_F(13, 2, 0.01)
    p = mf.default_params(reaction_policy=mf.MF_REACT_PER_FIBER)
    g = mf.BatchedFiberRVEAnalysisExplicit(nets, p)
    _, plan = g.plan()
    say(f"  plan {plan[:, :3].tolist()}")
    assert all(int(f) == 16384 for f in plan[:, 2]), plan
    s_g, st = g.run(F, return_status=True)
    assert (st == mf.MF_STATUS_OK).all()
    o = Oracle(nets, p)
    s_o = o.run(F)
    err = rel_err(s_g, s_o)
    say(f"  sigma rel err {err:.3e}")
    assert err <= FP64_TOL
    xo, uo, _ = o.get_state()
    xg, ug, _ = g.get_state()
    assert np.abs(xg - xo).max() <= 1e-14 and np.abs(ug - uo).max() <= 1e-14
    # forward tangent (chained trials) as well
    C_g = g.computeMaterialStiffness()
    assert rel_err(C_g, o.computeMaterialStiffness()) <= 1e-8


def test_relax_multi_rve_broadcast_trial(net27):
    say("case: relax, 3 RVEs, trial with broadcast F")
    nets = [net27, mf.generate_network(seed=5, n_nodes=300, E=1e8, radius=5e-5),
            mf.generate_network(seed=6, n_nodes=700, E=1e8, radius=5e-5)]
    F = strain_F(3, 3, 0.01)
    o = Oracle(nets)
    g = mf.BatchedFiberRVEAnalysisExplicit(nets)
    assert rel_err(g.run(F), o.run(F)) <= FP64_TOL
    Ft = np.diag([1.001, 0.999, 1.0])
    assert rel_err(g.run(Ft, update_coords=False), o.run(Ft, update_coords=False)) <= FP64_TOL
    xo, uo, _ = o.get_state()
    xg, ug, _ = g.get_state()
    assert np.array_equal(ug, uo)
    assert np.abs(xg - xo).max() <= 1e-14
    # broadcast F with update_coords mirrors the reference's extent check
    with pytest.raises(mf.MfError):
        g.run(Ft, update_coords=True)


# ---------------------------------------------------------------- ramp mode
def test_ramp_500_nodes_5000_steps_fp64_and_fp32c():
    """One RVE of 500 nodes, F = diag(1.05,1,1) ramped over 5000 steps."""
    say("case: one 500-node RVE, ramp, 5000 steps")
    net = mf.generate_network(seed=1, n_nodes=500)
    F = np.diag([1.05, 1.0, 1.0])
    s_o = Oracle([net], ramp_params()).run(F)
    g = mf.BatchedFiberRVEAnalysisExplicit([net], ramp_params())
    s_g = g.run(F)
    e64 = rel_err(s_g, s_o)
    g32 = mf.BatchedFiberRVEAnalysisExplicit([net], ramp_params(precision=mf.MF_PREC_FP32C))
    s32, st32 = g32.run(F, return_status=True)
    assert (st32 == 0).all()
    e32 = rel_err(s32, s_o)
    say(f"  fp64 {e64:.3e}  fp32c {e32:.3e}")
    assert e64 <= FP64_TOL
    assert e32 <= FP32C_TOL


def test_batch_500_nodes_subset_fp64_and_fp32c():
    """The 4096 x 500-node batch's generator and F distribution on 32 RVEs
    (seeds 1..32)."""
    R = 32
    say(f"case: 500-node batch subset R={R}")
    nets = [mf.generate_network(seed=1 + r, n_nodes=500) for r in range(R)]
    F = strain_F(2, R, 0.05)
    s_o = Oracle(nets, ramp_params()).run(F)
    g = mf.BatchedFiberRVEAnalysisExplicit(nets, ramp_params())
    s_g = g.run(F)
    e64 = rel_err(s_g, s_o)
    g32 = mf.BatchedFiberRVEAnalysisExplicit(nets, ramp_params(precision=mf.MF_PREC_FP32C))
    s32, st32 = g32.run(F, return_status=True)
    assert (st32 == 0).all()
    e32 = rel_err(s32, s_o)
    say(f"  fp64 {e64:.3e}  fp32c {e32:.3e}")
    assert e64 <= FP64_TOL
    assert e32 <= FP32C_TOL
    # the generated batch (mf_batch_create_generated) is the same batch
    gen = mf.default_gen_params(seed=1, n_nodes=500)
    g2 = mf.BatchedFiberRVEAnalysisExplicit(params=ramp_params(), generated=(gen, R))
    assert np.array_equal(g2.run(F), s_g)


def test_full_size_batch_4096x500_fp32c_and_fp64():
    """The full 4096 x 500-node, 5000-step batch: every FP32c stress within
    1e-4 of the FP64 GPU stress, FP64 equal to the oracle on RVEs sampled
    across the batch, every status OK, and a second FP32c batch built
    the same way giving bitwise the same stress: the kernels have no
    atomics or order-dependent sums."""
    R = 4096
    say(f"case: full 500-node batch R={R}, 5000 steps")
    gen = mf.default_gen_params(seed=1, n_nodes=500)
    F = strain_F(2, R, 0.05)
    g64 = mf.BatchedFiberRVEAnalysisExplicit(params=ramp_params(), generated=(gen, R))
    s64, st64 = g64.run(F, return_status=True)
    g32 = mf.BatchedFiberRVEAnalysisExplicit(
        params=ramp_params(precision=mf.MF_PREC_FP32C), generated=(gen, R))
    s32, st32 = g32.run(F, return_status=True)
    assert (st64 == 0).all() and (st32 == 0).all()
    assert np.isfinite(s32).all()
    e32 = rel_err(s32, s64)
    sample = [0, 1, 1365, 2730, 4094, 4095]
    nets = [mf.generate_network(seed=1 + r, n_nodes=500) for r in sample]
    s_o = Oracle(nets, ramp_params()).run(F[sample])
    e64 = rel_err(s64[sample], s_o)
    say(f"  fp32c vs fp64 max {e32:.3e}; fp64 vs oracle on {len(sample)} RVEs {e64:.3e}")
    assert e32 <= FP32C_TOL
    assert e64 <= FP64_TOL
    g32b = mf.BatchedFiberRVEAnalysisExplicit(
        params=ramp_params(precision=mf.MF_PREC_FP32C), generated=(gen, R))
    assert np.array_equal(g32b.run(F), s32)


def test_full_size_mixed_batch_16384_fp32c():
    """16384 RVEs of 200-4000 nodes (log-uniform), half the fibers on the
    toe law, F = I + sym(U) with U ~ U(-0.1, 0.1), 10000 ramp steps: every
    RVE of the one FP32c batch (all kernel-variant groups launched together)
    finishes with status OK and finite stress.  A sample goes to the oracle
    at the full 10000 steps: 128 RVEs spread over the batch, the 4 largest
    RVEs, and the largest RVE of every FP32c kernel variant in the plan.  On
    it the FP64 GPU solver matches the oracle to 1e-10 and the FP32c batch
    matches the oracle (and the FP64 GPU solver) to 1e-4."""
    R, S = 16384, 10000
    say(f"case: full mixed batch R={R}, 200-4000 nodes, {S} steps")
    u = mf.uniform(3, 11, R)
    sizes = np.rint(np.exp(np.log(200) + u * (np.log(4000) - np.log(200)))).astype(np.int32)
    F = strain_F(3, R, 0.1)
    gen = mf.default_gen_params(seed=100001, nonlinear_fraction=0.5)
    base = dict(n_steps=S, reaction_policy=mf.MF_REACT_PER_FIBER)
    g32 = mf.BatchedFiberRVEAnalysisExplicit(
        params=ramp_params(precision=mf.MF_PREC_FP32C, **base),
        generated=(gen, R), n_nodes=sizes)
    s32, st32 = g32.run(F, return_status=True)
    assert (st32 == 0).all()
    assert np.isfinite(s32).all()
    _, plan32 = g32.plan()
    g32.close()
    sub = set(int(r) for r in np.arange(0, R, R // 128)[:128])
    sub |= set(int(r) for r in np.argsort(sizes, kind="stable")[-4:])
    fam = {}
    for r in range(R):
        k = tuple(int(v) for v in plan32[r, :3])
        if k not in fam or sizes[r] > sizes[fam[k]]:
            fam[k] = r
    sub |= set(fam.values())
    sub = np.array(sorted(sub))
    say(f"  oracle sample: {len(sub)} RVEs ({len(fam)} FP32c variants), "
        f"{sizes[sub].min()}-{sizes[sub].max()} nodes")
    nets = [mf.generate_network(seed=100001 + int(r), n_nodes=int(sizes[r]),
                                nonlinear_fraction=0.5) for r in sub]
    g64 = mf.BatchedFiberRVEAnalysisExplicit(nets, ramp_params(**base))
    s64, st64 = g64.run(F[sub], return_status=True)
    assert (st64 == 0).all()
    s_o, st_o = Oracle(nets, ramp_params(**base)).run(F[sub], return_status=True)
    assert (st_o == 0).all()
    e64 = rel_err(s64, s_o)
    e32o = rel_err(s32[sub], s_o)
    e32 = rel_err(s32[sub], s64)
    big = np.argsort(sizes[sub], kind="stable")[-4:]
    say(f"  fp64 vs oracle {e64:.3e}; fp32c vs oracle {e32o:.3e}; fp32c vs fp64 {e32:.3e}; "
        f"4 largest ({sizes[sub][big].tolist()} nodes): fp64 {rel_err(s64[big], s_o[big]):.3e} "
        f"fp32c {rel_err(s32[sub][big], s_o[big]):.3e}")
    assert e64 <= FP64_TOL
    assert e32o <= FP32C_TOL
    assert e32 <= FP32C_TOL


def test_uneven_mixed_reactions():
    """The mixed 200-4000-node batch on a few RVEs: spread sizes, half the
    fibers on the toe law."""
    sizes = [200, 430, 910, 1650, 2700, 4000]
    say(f"case: uneven mixed sizes {sizes}")
    nets = [mf.generate_network(seed=100 + i, n_nodes=n, nonlinear_fraction=0.5)
            for i, n in enumerate(sizes)]
    F = strain_F(3, len(sizes), 0.1)
    p = dict(reaction_policy=mf.MF_REACT_PER_FIBER, n_steps=2000)
    s_o = Oracle(nets, ramp_params(**p)).run(F)
    g = mf.BatchedFiberRVEAnalysisExplicit(nets, ramp_params(**p))
    ng, plan = g.plan()
    say(f"  plan groups {ng}: {plan.tolist()}")
    s_g, st = g.run(F, return_status=True)
    assert (st == 0).all()
    e64 = rel_err(s_g, s_o)
    g32 = mf.BatchedFiberRVEAnalysisExplicit(nets, ramp_params(precision=mf.MF_PREC_FP32C, **p))
    s32, st32 = g32.run(F, return_status=True)
    assert (st32 == 0).all()
    e32 = rel_err(s32, s_o)
    say(f"  fp64 {e64:.3e}  fp32c {e32:.3e}")
    assert e64 <= FP64_TOL
    assert e32 <= FP32C_TOL


@pytest.mark.parametrize("block", [256, 512, 1024])
def test_kernel_variants_agree_bitwise(block):
    """Every thread count / LDS placement gives the same FP64 result."""
    say(f"case: variants block={block}")
    nets = [mf.generate_network(seed=40 + i, n_nodes=n) for i, n in enumerate([150, 700])]
    F = strain_F(4, 2, 0.05)
    ref = mf.BatchedFiberRVEAnalysisExplicit(nets, ramp_params(n_steps=800)).run(F)
    g = mf.BatchedFiberRVEAnalysisExplicit(nets, ramp_params(n_steps=800, block_size=block))
    assert np.array_equal(g.run(F), ref)


@pytest.mark.parametrize("scheme", [mf.MF_TANGENT_FORWARD, mf.MF_TANGENT_CENTRAL])
def test_ramp_tangent_fp64(scheme):
    """The 1000-2000-node tangent batch's shape: base + 6 forward or 12 central perturbations
    (h = 1e-6) in one launch (the fused run + tangent), then the same tangent
    from a separate run and tangent call."""
    R = 4
    say(f"case: ramp tangent scheme={scheme}")
    nets = [mf.generate_network(seed=300 + r, n_nodes=n, nonlinear_fraction=0.5)
            for r, n in enumerate([300, 500, 800, 1000])]
    F = strain_F(5, R, 0.05)
    p = ramp_params(n_steps=1500, reaction_policy=mf.MF_REACT_PER_FIBER,
                    tangent_scheme=scheme, fd_step=1e-6)
    o = Oracle(nets, p)
    s_o = o.run(F)
    C_o = o.computeMaterialStiffness()
    g = mf.BatchedFiberRVEAnalysisExplicit(nets, p)
    s_g, C_g, st = g.run_with_tangent(F, return_status=True)
    assert (st == mf.MF_STATUS_OK).all()
    es, ec = rel_err(s_g, s_o), rel_err(C_g, C_o)
    say(f"  sigma {es:.3e}  C {ec:.3e}")
    assert es <= FP64_TOL
    assert ec <= 1e-6
    # separate run + tangent gives the same numbers as the fused launch
    g2 = mf.BatchedFiberRVEAnalysisExplicit(nets, p)
    g2.run(F)
    C2, st2 = g2.computeMaterialStiffness(return_status=True)
    assert (st2 == mf.MF_STATUS_OK).all()
    assert np.array_equal(C2, C_g)


def _affine_patch():
    rng = np.random.default_rng(3)
    X = rng.uniform(-0.5, 0.5, (12, 3))
    fb = np.array([[i, j] for i in range(12) for j in range(i + 1, 12) if (i + j) % 3 == 0],
                  np.int32)
    return mf.Network(X, fb, np.arange(12, dtype=np.int32), [mf.Reaction(0, 0.01, 1e5, 1000.0)])


def _perturbation(i, h):
    """Almansi perturbation of Voigt direction i (BatchedFiberRVEAnalysisExplicit.h:338-390)."""
    Fp = np.eye(3)
    if i < 3:
        Fp[i, i] = math.sqrt(1.0 / (1 - 2 * h))
        return Fp
    l2, l3 = (1 + h) / (1 - h * h), (1 - h) / (1 - h * h)
    j, k = [(1, 2), (0, 2), (0, 1)][i - 3]
    Fp[j, j] = Fp[k, k] = 0.5 * (math.sqrt(l2) + math.sqrt(l3))
    Fp[j, k] = Fp[k, j] = 0.5 * (math.sqrt(l2) - math.sqrt(l3))
    return Fp


@pytest.mark.parametrize("scheme", [mf.MF_TANGENT_FORWARD, mf.MF_TANGENT_CENTRAL])
def test_ramp_tangent_affine_patch_closed_form(scheme):
    """The GPU ramp-mode tangent on the all-boundary patch after two update
    runs F1, F2 equals the difference quotient of the closed-form affine
    stresses at P_i(+-h) F2 F1 X over det F1 det F2, in both precisions."""
    say(f"case: ramp tangent, all-boundary patch, scheme={scheme}")
    net = _affine_patch()
    X, fb = net.coords, net.fibers
    r0 = net.reactions[0]
    EA = r0.E * math.pi * r0.radius ** 2
    F1 = np.array([[1.02, 0.01, 0.0], [0.0, 0.99, 0.005], [0.003, 0.0, 1.01]])
    F2 = np.array([[0.995, 0.0, 0.012], [0.004, 1.015, 0.0], [0.0, -0.008, 1.0]])
    V = np.linalg.det(F1) * np.linalg.det(F2)
    h = 1e-6
    central = scheme == mf.MF_TANGENT_CENTRAL

    def sig(x):
        d = x[fb[:, 1]] - x[fb[:, 0]]
        l = np.linalg.norm(d, axis=1)
        l0 = np.linalg.norm(X[fb[:, 1]] - X[fb[:, 0]], axis=1)
        lam = l / l0
        S = np.einsum("e,ei,ej->ij", lam * EA * 0.5 * (lam ** 2 - 1) / l, d, d) / V
        return np.array([S[0, 0], S[1, 1], S[2, 2], S[1, 2], S[0, 2], S[0, 1]])

    x1 = X @ F1.T
    s0 = sig(x1 @ F2.T)
    Cc = np.zeros((6, 6))
    for i in range(6):
        t = [sig(x1 @ (_perturbation(i, sg * h) @ F2).T) for sg in ((1, -1) if central else (1,))]
        Cc[:, i] = (t[0] - t[1]) / (2 * h) if central else (t[0] - s0) / h
    for prec in (mf.MF_PREC_FP64, mf.MF_PREC_FP32C):
        p = ramp_params(n_steps=200, tangent_scheme=scheme, fd_step=h, precision=prec,
                        tangent_precision=mf.MF_PREC_FP64)
        g = mf.BatchedFiberRVEAnalysisExplicit([net], p)
        g.run(F1)
        s_c = g.run(F2)[0]
        C, st = g.computeMaterialStiffness(return_status=True)
        e_s = np.abs(s_c - s0).max() / np.abs(s0).max()
        e_c = np.abs(C[0] - Cc).max() / np.abs(Cc).max()
        say(f"  precision {prec}: sigma {e_s:.2e}  C {e_c:.2e}")
        assert (st == 0).all()
        assert e_s <= (1e-12 if prec == mf.MF_PREC_FP64 else 1e-6)
        assert e_c <= 1e-7


@pytest.mark.parametrize("mode", [mf.MF_MODE_RAMP, mf.MF_MODE_RELAX])
def test_forward_tangent_mixed_precision(mode, golden, net27):
    """run() in FP32c with the forward tangent's solves in FP64: the base of
    the differences is an FP64 solve too, so C matches the all-FP64 tangent
    instead of carrying (sigma_fp32c - sigma_fp64)/h.  Ramp mode: the same
    FP64 solves from the same state, equal to rounding of the stress sums.
    Relax mode (Network_2_7 under the golden F, reference settings): equal to
    the relaxation tolerance the reference's own golden tangent blocks show
    (5e-3)."""
    say(f"case: forward tangent, FP32c run, FP64 tangent solves, mode={mode}")
    if mode == mf.MF_MODE_RAMP:
        nets = [mf.generate_network(seed=330 + r, n_nodes=n, E=1e8, radius=5e-5,
                                    nonlinear_fraction=0.5) for r, n in enumerate([300, 600])]
        F = strain_F(7, 2, 0.02)
        kw = dict(mode=mode, n_steps=1500, reaction_policy=mf.MF_REACT_PER_FIBER, fd_step=1e-6)
    else:
        nets = [net27]
        F = np.array(golden["network_2_7"]["F"])[None]
        kw = dict(mode=mode, max_steps=200000)
    kw.update(tangent_scheme=mf.MF_TANGENT_FORWARD, tangent_precision=mf.MF_PREC_FP64)
    g64 = mf.BatchedFiberRVEAnalysisExplicit(nets, mf.default_params(**kw))
    s64 = g64.run(F)
    C64 = g64.computeMaterialStiffness()
    gm = mf.BatchedFiberRVEAnalysisExplicit(
        nets, mf.default_params(precision=mf.MF_PREC_FP32C, **kw))
    s32 = gm.run(F)
    Cm, st = gm.computeMaterialStiffness(return_status=True)
    assert (st == mf.MF_STATUS_OK).all()
    e = rel_err(Cm, C64)
    say(f"  sigma fp32c vs fp64 {rel_err(s32, s64):.2e}; C mixed vs fp64 {e:.2e}")
    assert e <= (1e-9 if mode == mf.MF_MODE_RAMP else 5e-3)
    # the fused entry point takes the same route
    gf = mf.BatchedFiberRVEAnalysisExplicit(
        nets, mf.default_params(precision=mf.MF_PREC_FP32C, **kw))
    _, Cf = gf.run_with_tangent(F)
    assert rel_err(Cf, C64) <= (1e-9 if mode == mf.MF_MODE_RAMP else 5e-3)


def test_energies_after_fused_run_with_tangent():
    """The fused ramp run + tangent tracks no energies, so the energies of
    an earlier run are not reported as its own."""
    say("case: energies after run_with_tangent")
    nets = [mf.generate_network(seed=340, n_nodes=300)]
    p = ramp_params(n_steps=300, energy_diagnostics=1, tangent_scheme=mf.MF_TANGENT_CENTRAL)
    g = mf.BatchedFiberRVEAnalysisExplicit(nets, p)
    g.run(np.diag([1.01, 1.0, 1.0]))
    assert np.isfinite(g.last_energies()).all()
    g.run_with_tangent(np.diag([1.005, 1.0, 1.0])[None])
    with pytest.raises(mf.MfError):
        g.last_energies()


def test_full_size_central_tangent_4096():
    """4096 RVEs of 1000-2000 nodes, half the fibers on the toe law, base
    solve + 12 central perturbations (h = 1e-6) of 5000 ramp steps each in
    one FP64 launch: every status OK, every stress and tangent entry finite.
    A second tangent at h = 1e-5 from the same state ranks the RVEs by how
    much their tangent moves with h (fibers near a toe-law transition).
    Against the oracle (full 5000 steps, h = 1e-6): the smallest and largest
    RVEs, the 8 most h-sensitive ones and 14 spread over the batch, stress
    to 1e-10 and tangent to 1e-6."""
    R, S = 4096, 5000
    say(f"case: full central tangent R={R}, 1000-2000 nodes, {S} steps, 13 solves each")
    u = mf.uniform(4, 11, R)
    sizes = (1000 + np.floor(u * 1001)).astype(np.int32)
    F = strain_F(4, R, 0.05)
    gen = mf.default_gen_params(seed=200001, nonlinear_fraction=0.5)
    p = ramp_params(n_steps=S, reaction_policy=mf.MF_REACT_PER_FIBER,
                    tangent_scheme=mf.MF_TANGENT_CENTRAL, fd_step=1e-6)
    g = mf.BatchedFiberRVEAnalysisExplicit(params=p, generated=(gen, R), n_nodes=sizes)
    s_g, C_g, st = g.run_with_tangent(F, return_status=True)
    assert (st == mf.MF_STATUS_OK).all()
    assert np.isfinite(s_g).all() and np.isfinite(C_g).all()
    g.set_params(fd_step=1e-5)
    C5, st5 = g.computeMaterialStiffness(return_status=True)
    assert (st5 == mf.MF_STATUS_OK).all()
    g.close()
    sens = np.abs(C_g - C5).reshape(R, -1).max(1) / np.abs(C5).reshape(R, -1).max(1)
    pick = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    pick |= set(int(r) for r in np.argsort(sens, kind="stable")[-8:])
    pick |= set(int(r) for r in np.arange(0, R, R // 14)[:14])
    pick = np.array(sorted(pick))
    nets = [mf.generate_network(seed=200001 + int(r), n_nodes=int(sizes[r]),
                                nonlinear_fraction=0.5) for r in pick]
    o = Oracle(nets, p)
    s_o = o.run(F[pick])
    C_o = o.computeMaterialStiffness()
    es, ec = rel_err(s_g[pick], s_o), rel_err(C_g[pick], C_o)
    top = np.argsort(sens[pick], kind="stable")[-8:]
    say(f"  {len(pick)} RVEs ({sizes[pick].min()}-{sizes[pick].max()} nodes) vs oracle: "
        f"sigma {es:.3e}  C {ec:.3e}; 8 most h-sensitive (|C(1e-6)-C(1e-5)| up to "
        f"{sens.max():.2e}): C {rel_err(C_g[pick][top], C_o[top]):.3e}")
    assert es <= FP64_TOL
    assert ec <= 1e-6


def test_ramp_trial_restores_state():
    say("case: ramp trial run leaves state unchanged")
    net = mf.generate_network(seed=9, n_nodes=400)
    g = mf.BatchedFiberRVEAnalysisExplicit([net], ramp_params(n_steps=500))
    g.run(np.diag([1.01, 1, 1]))
    x0, u0, V0 = g.get_state()
    g.run(np.diag([1.02, 1, 1]), update_coords=False)
    x1, u1, V1 = g.get_state()
    assert np.array_equal(x0, x1) and np.array_equal(u0, u1) and np.array_equal(V0, V1)


def test_orientation_tensors_match_oracle(golden, net27):
    say("case: orientation tensors")
    F = np.array(golden["network_2_7"]["F"])
    o = Oracle([net27])
    g = mf.BatchedFiberRVEAnalysisExplicit([net27])
    om = g.compute3DOrientationTensor()
    assert np.allclose(om[0], golden["network_2_7"]["summary_orientation"], atol=1.5e-6)
    o.run(F)
    g.run(F)
    assert rel_err(g.compute3DOrientationTensor(), o.orientation()) <= 1e-13
    n = [0.3, 0.2, 0.9]
    assert rel_err(g.compute2DOrientationTensor(n), o.orientation(n)) <= 1e-13
    # non-unit normals: the reference divides by |n|^2, a quirk both keep
    for n in ([0.0, 0.0, 2.0], [0.0, 0.5, 0.0], [1.5, -0.7, 0.4]):
        assert rel_err(g.compute2DOrientationTensor(n), o.orientation(n)) <= 1e-13


def test_edge_cases_all_boundary_and_no_boundary():
    say("case: edge cases")
    rng = np.random.default_rng(3)
    X = rng.uniform(-0.5, 0.5, (12, 3))
    fb = np.array([[i, j] for i in range(12) for j in range(i + 1, 12) if (i + j) % 3 == 0],
                  np.int32)
    r = [mf.Reaction(0, 0.01, 1e5, 1000.0)]
    all_b = mf.Network(X, fb, np.arange(12, dtype=np.int32), r)
    no_b = mf.Network(X, fb, np.zeros(0, np.int32), r)
    F = np.array([[1.02, 0.01, 0.0], [0.0, 0.99, 0.005], [0.003, 0.0, 1.01]])
    for mode in (mf.MF_MODE_RELAX, mf.MF_MODE_RAMP):
        p = mf.default_params(mode=mode, n_steps=300)
        nets = [all_b, no_b]
        s_o = Oracle(nets, p).run(np.stack([F, F]))
        s_g = mf.BatchedFiberRVEAnalysisExplicit(nets, p).run(np.stack([F, F]))
        assert rel_err(s_g[:1], s_o[:1]) <= FP64_TOL
        assert np.all(s_g[1] == 0) and np.all(s_o[1] == 0)
        # the FP32c families with no free node and with no boundary node
        for fk in (mf.MF_FP32_KERNEL_NODE, mf.MF_FP32_KERNEL_FIBER):
            p32 = mf.default_params(mode=mode, n_steps=300, precision=mf.MF_PREC_FP32C,
                                    fp32_kernel=fk)
            s32 = mf.BatchedFiberRVEAnalysisExplicit(nets, p32).run(np.stack([F, F]))
            assert rel_err(s32[:1], s_o[:1]) <= FP32C_TOL
            assert np.all(s32[1] == 0)


def test_status_nan_and_step_cap():
    say("case: status codes")
    X = np.array([[-0.5, 0, 0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.5, 0, 0]])
    fb = np.array([[0, 1], [1, 2], [2, 3]], np.int32)
    r = [mf.Reaction(0, 0.01, 1e6, 1e3)]
    deg = mf.Network(X, fb, np.array([0, 3], np.int32), r)
    g = mf.BatchedFiberRVEAnalysisExplicit([deg])
    _, st = g.run(np.diag([1.01, 1, 1]), return_status=True)
    assert st[0] == mf.MF_STATUS_NAN
    net = mf.generate_network(seed=2, n_nodes=200, E=1e8, radius=5e-5)
    g = mf.BatchedFiberRVEAnalysisExplicit([net], mf.default_params(tolerance=1e-30,
                                                                    max_steps=300))
    _, st = g.run(np.diag([1.01, 1, 1]), return_status=True)
    assert st[0] == mf.MF_STATUS_STEP_CAP


def test_deterministic_repeat():
    say("case: determinism")
    nets = [mf.generate_network(seed=70 + i, n_nodes=500) for i in range(8)]
    F = strain_F(6, 8, 0.05)
    for prec in (mf.MF_PREC_FP64, mf.MF_PREC_FP32C):
        a = mf.BatchedFiberRVEAnalysisExplicit(nets, ramp_params(n_steps=1000, precision=prec))
        b = mf.BatchedFiberRVEAnalysisExplicit(nets, ramp_params(n_steps=1000, precision=prec))
        assert np.array_equal(a.run(F), b.run(F))


@pytest.mark.parametrize("nl", [0.0, 0.5])
def test_one_rve_per_wave_agrees_with_one_rve_per_workgroup(nl):
    """Small RVEs (120-180 nodes) on the one-RVE-per-wave register kernel,
    one and four RVEs per workgroup (seven RVEs: the last workgroup of four
    has a spare wave), against the one-RVE-per-workgroup register kernel.
    Ramp mode: the same per-node arithmetic gives bitwise the same committed
    state, the stress differs only by the order of its FP64 boundary sum,
    and both are within the FP32c bar of the oracle.  Relax mode: every
    solve converges and the two agree to 1e-3 (the residual's FP64 sum order
    moves the stopping step, and at the default tolerance the stress still
    drifts by ~1e-4 between nearby steps)."""
    say(f"case: one RVE per wave, nonlinear={nl}")
    for mode in (mf.MF_MODE_RAMP, mf.MF_MODE_RELAX):
        extra = {} if mode == mf.MF_MODE_RAMP else dict(E=1e8, radius=5e-5)
        nets = [mf.generate_network(seed=300 + i, n_nodes=n, nonlinear_fraction=nl, **extra)
                for i, n in enumerate([120, 130, 140, 150, 160, 170, 180])]
        F = strain_F(15, len(nets), 0.06 if mode == mf.MF_MODE_RAMP else 0.01)
        common = dict(mode=mode, n_steps=1500, precision=mf.MF_PREC_FP32C,
                      reaction_policy=mf.MF_REACT_PER_FIBER, chunks=1)
        out = {}
        for fk, bs in ((mf.MF_FP32_KERNEL_NODE, 256), (mf.MF_FP32_KERNEL_WAVE, 64),
                       (mf.MF_FP32_KERNEL_WAVE, 256)):
            g = mf.BatchedFiberRVEAnalysisExplicit(
                nets, mf.default_params(fp32_kernel=fk, block_size=bs, **common))
            _, plan = g.plan()
            wave = [bool(int(f) & 4096) for f in plan[:, 2]]
            assert all(wave) == (fk == mf.MF_FP32_KERNEL_WAVE) and any(wave) == all(wave), plan
            s, st = g.run(F, return_status=True)
            assert (st == mf.MF_STATUS_OK).all()
            out[(fk, bs)] = (s, g.get_state()[1])
        ref = out[(mf.MF_FP32_KERNEL_NODE, 256)]
        w1, w4 = out[(mf.MF_FP32_KERNEL_WAVE, 64)], out[(mf.MF_FP32_KERNEL_WAVE, 256)]
        assert np.array_equal(w1[0], w4[0]) and np.array_equal(w1[1], w4[1])
        e_ref = rel_err(w1[0], ref[0])
        say(f"  mode {mode}: wave vs workgroup stress {e_ref:.2e}, "
            f"state bitwise {np.array_equal(w1[1], ref[1])}")
        if mode == mf.MF_MODE_RAMP:
            assert np.array_equal(w1[1], ref[1])
            assert e_ref <= 1e-13
            e_o = rel_err(w1[0], Oracle(nets, mf.default_params(
                **dict(common, precision=mf.MF_PREC_FP64))).run(F))
            say(f"  wave vs oracle {e_o:.2e}")
            assert e_o <= FP32C_TOL
        else:
            assert e_ref <= 1e-3


def test_register_kernel_slots_agree_bitwise():
    """The register-resident FP32c kernel with one or several free nodes per
    thread gives bitwise the same state (per-node arithmetic is the same),
    and stays within the FP32c bar of the oracle."""
    say("case: register kernel, 1-4 nodes per thread")
    for nl in (0.0, 0.5):
        nets = [mf.generate_network(seed=60 + i, n_nodes=n, nonlinear_fraction=nl)
                for i, n in enumerate([700, 1000])]
        F = strain_F(6, 2, 0.05)
        node = dict(n_steps=600, precision=mf.MF_PREC_FP32C,
                    reaction_policy=mf.MF_REACT_PER_FIBER, fp32_kernel=mf.MF_FP32_KERNEL_NODE)
        p = ramp_params(**node)
        g = mf.BatchedFiberRVEAnalysisExplicit(nets, p)
        _, plan = g.plan()
        ref = g.run(F)
        assert np.all(plan[:, 1] == 1), plan  # one node per thread by default
        assert np.all(plan[:, 2] & 4), plan  # register-resident lists
        forced = mf.BatchedFiberRVEAnalysisExplicit(nets, ramp_params(block_size=512, **node))
        _, fplan = forced.plan()
        say(f"  nonlinear {nl}: default {plan[:, :2].tolist()} forced {fplan[:, :2].tolist()}")
        assert np.all(fplan[:, 1] >= 2) and np.all(fplan[:, 2] & 4), fplan
        # per-node arithmetic is the same: the committed state agrees bitwise;
        # the FP64 stress sum runs over a different number of threads
        sf = forced.run(F)
        ref_x, ref_u, _ = g.get_state()
        fx, fu, _ = forced.get_state()
        assert np.array_equal(fx, ref_x) and np.array_equal(fu, ref_u)
        assert rel_err(sf, ref) <= 1e-13
        o = Oracle(nets, p)
        err = rel_err(ref, o.run(F))
        say(f"  vs oracle {err:.2e}")
        assert err <= FP32C_TOL


def test_fiber_kernels_match_oracle():
    """The two fiber-evaluating FP32c kernel families, each fiber evaluated
    once per step: fiber-parallel (end forces into per-node slots; linear,
    one-linear-one-nonlinear and reaction-table modes) and fiber-streamed
    (per-fiber force records, nodes gather by reference; large RVEs).  Within
    the FP32c bar of the oracle, and close to the node-centric kernels on the
    same solves (the families round the per-fiber contributions differently)."""
    say("case: fiber kernels, three reaction modes, 480-2400 nodes")
    nets = [mf.generate_network(seed=80 + i, n_nodes=n, nonlinear_fraction=nl)
            for i, (n, nl) in enumerate([(480, 0.0), (1150, 0.5), (1400, 0.5),
                                         (2400, 0.5), (2400, 0.0)])]
    # a second nonlinear reaction (B = 2) puts the RVE in reaction-table mode
    t = mf.generate_network(seed=90, n_nodes=480, nonlinear_fraction=0.5)
    t.reactions.append(dataclasses.replace(t.reactions[1], B=2.0))
    t.fiber_reaction = (np.arange(t.n_fibers) % 3).astype(np.int32)
    nets.append(t)
    F = strain_F(8, len(nets), 0.08)
    common = dict(n_steps=1500, reaction_policy=mf.MF_REACT_PER_FIBER)
    fib = mf.BatchedFiberRVEAnalysisExplicit(
        nets, ramp_params(precision=mf.MF_PREC_FP32C, fp32_kernel=mf.MF_FP32_KERNEL_FIBER, **common))
    _, plan = fib.plan()
    say(f"  fiber plan {plan[:, :3].tolist()}")
    fam = [int(f) & (64 | 128) for f in plan[:, 2]]
    assert all(fam), plan
    assert 64 in fam and 128 in fam, fam  # both families exercised
    modes = [int(f) & (8 | 32) for f in plan[:, 2]]
    assert modes == [8, 32, 32, 32, 8, 0], modes
    s_f, st = fib.run(F, return_status=True)
    assert (st == mf.MF_STATUS_OK).all()
    node = mf.BatchedFiberRVEAnalysisExplicit(
        nets, ramp_params(precision=mf.MF_PREC_FP32C, fp32_kernel=mf.MF_FP32_KERNEL_NODE, **common))
    _, nplan = node.plan()
    assert not np.any(nplan[:, 2] & (64 | 128)), nplan
    s_n = node.run(F)
    s_o = Oracle(nets, ramp_params(**common)).run(F)
    e_f, e_n, e_fn = rel_err(s_f, s_o), rel_err(s_n, s_o), rel_err(s_f, s_n)
    say(f"  fiber vs oracle {e_f:.2e}, node vs oracle {e_n:.2e}, fiber vs node {e_fn:.2e}")
    assert e_f <= FP32C_TOL and e_n <= FP32C_TOL
    # committed displacements of the two families agree to FP32 accumulation
    _, u_f, _ = fib.get_state()
    _, u_n, _ = node.get_state()
    assert np.abs(u_f - u_n).max() <= 1e-4 * np.abs(u_n).max()


def test_fiber_streamed_variants_agree_bitwise():
    """The fiber-streamed FP32c kernel at 768 threads (each node's first
    force-record references held in registers) and at 1024 threads
    (references streamed from global memory) evaluates every fiber the same
    way and sums each node's records in the same list order: bitwise the
    same stress and state, within the FP32c bar of the oracle.  Toe-law and
    linear RVEs of 1800-2400 nodes (4400-5900 fibers: more than the
    fiber-parallel shapes hold; 3-4 nodes and 6-8 fibers per thread)."""
    say("case: fiber-streamed kernel, 768 vs 1024 threads")
    nets = [mf.generate_network(seed=95 + i, n_nodes=n, nonlinear_fraction=nl)
            for i, (n, nl) in enumerate([(1800, 0.5), (2400, 0.5), (2400, 0.0)])]
    F = strain_F(12, len(nets), 0.08)
    common = dict(n_steps=1000, reaction_policy=mf.MF_REACT_PER_FIBER)
    out = {}
    for block in (768, 1024):
        g = mf.BatchedFiberRVEAnalysisExplicit(
            nets, ramp_params(precision=mf.MF_PREC_FP32C, fp32_kernel=mf.MF_FP32_KERNEL_FIBER,
                              block_size=block, **common))
        _, plan = g.plan()
        say(f"  {block} threads: plan {plan[:, :3].tolist()}")
        assert all(int(f) & 128 for f in plan[:, 2]), plan  # fiber-streamed
        s, st = g.run(F, return_status=True)
        assert (st == mf.MF_STATUS_OK).all()
        out[block] = (s, g.get_state()[1])
    assert np.array_equal(out[768][0], out[1024][0])
    assert np.array_equal(out[768][1], out[1024][1])
    e = rel_err(out[768][0], Oracle(nets, ramp_params(**common)).run(F))
    say(f"  vs oracle {e:.2e}")
    assert e <= FP32C_TOL


def test_large_rve_global_state_kernel():
    """An RVE with more free nodes (~5400) than the register and LDS kernels
    have lanes takes the FP64 kernel with the state in global memory: it
    matches the oracle like the other FP64 kernels, and in an FP32c batch the
    same RVE falls back to it (bitwise the FP64 result) while a small RVE
    keeps an FP32c kernel."""
    say("case: 6000-node RVE, global-state FP64 kernel")
    big = mf.generate_network(seed=7, n_nodes=6000, nonlinear_fraction=0.5)
    small = mf.generate_network(seed=8, n_nodes=400, nonlinear_fraction=0.5)
    common = dict(n_steps=400, reaction_policy=mf.MF_REACT_PER_FIBER)
    F = strain_F(12, 2, 0.05)
    g64 = mf.BatchedFiberRVEAnalysisExplicit([big, small], ramp_params(**common))
    _, p64 = g64.plan()
    say(f"  FP64 plan {p64[:, :3].tolist()}")
    assert p64[0, 2] == 8192 and p64[1, 2] != 8192, p64
    s64, st = g64.run(F, return_status=True)
    assert (st == mf.MF_STATUS_OK).all()
    s_o = Oracle([big, small], ramp_params(**common)).run(F)
    err = rel_err(s64, s_o)
    say(f"  FP64 vs oracle {err:.2e}")
    assert err <= FP64_TOL
    x64, u64, _ = g64.get_state()
    g32 = mf.BatchedFiberRVEAnalysisExplicit([big, small],
                                             ramp_params(precision=mf.MF_PREC_FP32C, **common))
    _, p32 = g32.plan()
    say(f"  FP32c plan {p32[:, :3].tolist()}")
    assert p32[0, 2] == 8192 and p32[1, 2] != 8192, p32
    s32 = g32.run(F)
    assert np.array_equal(s32[0], s64[0])
    assert rel_err(s32[1:], s_o[1:]) <= FP32C_TOL
    x32, u32, _ = g32.get_state()
    n0 = big.n_nodes
    assert np.array_equal(u32[:n0], u64[:n0])
    # hipGraph replays of trial runs equal the eager trial run: the
    # global-state scratch exists before the capture
    Fd = mf.DeviceArray.from_host(F.reshape(2, 9))
    sd = mf.DeviceArray((2, 6))
    ge = mf.BatchedFiberRVEAnalysisExplicit([big, small], ramp_params(**common))
    ge.run_device(Fd, sd)
    ge.synchronize()
    s_eager = sd.numpy().copy()
    gg = mf.BatchedFiberRVEAnalysisExplicit([big, small], ramp_params(**common))
    gg.set_graph(True)
    for _ in range(3):
        gg.run_device(Fd, sd)
        gg.synchronize()
        assert np.array_equal(sd.numpy(), s_eager)


def test_reference_network_del_4450_global_state():
    """The reference's largest test network, del_4450seedL5_new_1 (12308
    nodes, 38376 fibers; test/microscale/fiber_networks, read by the
    reference's global-stiffness test), through the C ABI.  It spans
    [-2.5, 2.5]^3, so no node lies on the faces of the [-0.5, 0.5]^3 RVE
    (RVE.h:86-103) and every node is free: the stress is zero and the
    relax loop moves all 12308 nodes, beyond every LDS kernel, on the FP64
    global-state kernel.  300 steps of the reference loop (tolerance out of
    reach) leave the state bitwise as in the oracle."""
    say("case: del_4450seedL5_new_1, 12308 free nodes")
    net = mf.load_network(os.path.join(GOLDEN, "del_4450", "del_4450seedL5_new_1.txt"))
    assert net.n_nodes == 12308 and net.n_fibers == 38376 and len(net.boundary) == 0
    F = np.array([[1.01, 0.002, 0.0], [0.0, 0.995, 0.001], [0.0, 0.0, 1.0]])
    p = mf.default_params(tolerance=1e-30, max_steps=300)
    g = mf.BatchedFiberRVEAnalysisExplicit([net], p)
    _, plan = g.plan()
    assert plan[0, 2] == 8192, plan
    s_g, st = g.run(F, return_status=True)
    o = Oracle([net], p)
    s_o, st_o = o.run(F, return_status=True)
    assert st[0] == st_o[0] == mf.MF_STATUS_STEP_CAP
    assert np.all(s_g == 0) and np.all(s_o == 0)
    xg, ug, _ = g.get_state()
    xo, uo, _ = o.get_state()
    say(f"  max|u| {np.abs(uo).max():.3e}, max|du| {np.abs(ug - uo).max():.3e}")
    assert np.array_equal(ug, uo) and np.array_equal(xg, xo)


def test_fp32c_families_relax_mode_and_step_cap():
    """Reference (relax) mode through each FP32c kernel family: node-centric
    (register lists), fiber-parallel (300 nodes) and fiber-streamed (1300
    nodes, toe law), against the FP64 GPU solve (which the other tests pin to
    the oracle).  With tolerance 1e-6 every family converges (status OK); the
    stress there still drifts by ~1e-3 between 1e-6 and full convergence, so
    solvers that stop at different dt refreshes agree to that level only.
    With the tolerance out of reach, every family stops at the step cap (same
    step as FP64) and matches the FP64 stress within the FP32c bar."""
    say("case: FP32c families in relax mode")
    nets = [mf.generate_network(seed=21, n_nodes=300, E=1e8, radius=5e-5),
            mf.generate_network(seed=22, n_nodes=1300, E=1e8, radius=5e-5,
                                nonlinear_fraction=0.5)]
    F = strain_F(9, len(nets), 0.01)
    conv = dict(reaction_policy=mf.MF_REACT_PER_FIBER, tolerance=1e-6)
    capped = dict(reaction_policy=mf.MF_REACT_PER_FIBER, tolerance=1e-30, max_steps=100000)
    s64 = mf.BatchedFiberRVEAnalysisExplicit(nets, mf.default_params(**conv)).run(F)
    c64 = mf.BatchedFiberRVEAnalysisExplicit(nets, mf.default_params(**capped)).run(F)
    for fk in (mf.MF_FP32_KERNEL_NODE, mf.MF_FP32_KERNEL_FIBER):
        g = mf.BatchedFiberRVEAnalysisExplicit(
            nets, mf.default_params(precision=mf.MF_PREC_FP32C, fp32_kernel=fk, **conv))
        _, plan = g.plan()
        s_g, st = g.run(F, return_status=True)
        g_c = mf.BatchedFiberRVEAnalysisExplicit(
            nets, mf.default_params(precision=mf.MF_PREC_FP32C, fp32_kernel=fk, **capped))
        c_g, st_c = g_c.run(F, return_status=True)
        e_conv, e_cap = rel_err(s_g, s64), rel_err(c_g, c64)
        say(f"  family {fk}: plan {plan[:, :3].tolist()}, converged {st.tolist()} "
            f"vs FP64 {e_conv:.2e}; capped {st_c.tolist()} vs FP64 {e_cap:.2e}")
        if fk == mf.MF_FP32_KERNEL_FIBER:
            assert plan[0, 2] & 64 and plan[1, 2] & 128, plan  # both fiber kernels
        assert (st == mf.MF_STATUS_OK).all() and e_conv <= 1e-3
        assert (st_c == mf.MF_STATUS_STEP_CAP).all() and e_cap <= FP32C_TOL


def test_energy_diagnostics_and_smoothstep_match_oracle(net27):
    """Energy accounting (FP64 energy variants) and the SmoothAmp ramp against
    the oracle; the balance check stops the same solves."""
    say("case: energy diagnostics + smoothstep amplitude")
    nets = [net27] + [mf.generate_network(seed=70 + i, n_nodes=n, nonlinear_fraction=0.5)
                      for i, n in enumerate([400, 900])]
    F = strain_F(7, 3, 0.04)
    for amp in (mf.MF_AMP_LINEAR, mf.MF_AMP_SMOOTHSTEP):
        p = ramp_params(n_steps=1500, amplitude=amp, energy_diagnostics=1,
                        reaction_policy=mf.MF_REACT_PER_FIBER)
        o = Oracle(nets, p)
        so = o.run(F)
        Wo = o.last_energies()
        g = mf.BatchedFiberRVEAnalysisExplicit(nets, p)
        sg, st = g.run(F, return_status=True)
        Wg = g.last_energies()
        es, ew = rel_err(sg, so), rel_err(Wg, Wo)
        say(f"  amplitude {amp}: sigma {es:.2e} energies {ew:.2e}  W0 {Wg[0].round(12).tolist()}")
        assert np.all(st == mf.MF_STATUS_OK)
        assert es <= FP64_TOL and ew <= 1e-10
        p32 = ramp_params(n_steps=1500, amplitude=amp, precision=mf.MF_PREC_FP32C,
                          reaction_policy=mf.MF_REACT_PER_FIBER)
        s32 = mf.BatchedFiberRVEAnalysisExplicit(nets, p32).run(F)
        assert rel_err(s32, so) <= FP32C_TOL
    # balance check: a tight eps stops every solve at its first check
    p = ramp_params(n_steps=1500, energy_diagnostics=1, energy_check_eps=1e-14,
                    reaction_policy=mf.MF_REACT_PER_FIBER)
    so, sto = Oracle(nets, p).run(F, return_status=True)
    sg, stg = mf.BatchedFiberRVEAnalysisExplicit(nets, p).run(F, return_status=True)
    assert np.array_equal(stg, sto) and np.all(stg == mf.MF_STATUS_ENERGY)
    assert rel_err(sg, so) <= FP64_TOL
    with pytest.raises(mf.MfError):
        mf.BatchedFiberRVEAnalysisExplicit(nets, ramp_params(energy_diagnostics=1,
                                                             precision=mf.MF_PREC_FP32C))
