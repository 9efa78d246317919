"""Pins the CPU FP64 oracle against the reference's own numbers and against
closed forms (CPU only)."""
import math

import numpy as np
import pytest

import mumfim_amd as mf
from conftest import voigt
from oracle.oracle import Oracle, fiber_force, reorder_permutation


def test_reorder_known_answer(golden):
    """ReorderMesh KAT of test/microscale/test_reorder_mesh.cc:9-66."""
    k = golden["reorder_kat"]
    perm = reorder_permutation(k["nverts"], k["fixed"])
    conn = [int(perm[c]) for c in k["initial_connectivity"]]
    assert conn == k["final_connectivity"]
    coords = np.arange(24, dtype=float).reshape(8, 3)
    out = np.zeros_like(coords)
    out[perm] = coords
    assert out.ravel().tolist() == [float(v) for v in k["final_coordinates"]]


def _golden_sigma(golden, block="explicit_cpu", i=0):
    g = golden["network_2_7"]
    F = np.array(g["F"])
    J = np.linalg.det(F)
    blk = g[block] if isinstance(g[block], dict) else g[block][i]
    return F, J, np.array(voigt(blk["stress33"])), np.array(blk["C"])


def test_network27_stress_matches_golden(golden, net27):
    """Batched stress = non-batched golden / det F
    (explicit_cpu_out.txt:28-31; the batched path divides by V)."""
    F, J, s_gold, _ = _golden_sigma(golden)
    o = Oracle([net27])
    s = o.run(F)[0]
    scale = np.abs(s_gold).max() / J
    # 6 printed significant digits: 5e-7 absolute on the largest entry
    assert np.abs(s - s_gold / J).max() <= 2e-6 * scale
    # the GPU-path output of the same solver carries the same digits
    _, _, s_gpu, _ = _golden_sigma(golden, "explicit_gpu")
    assert np.abs(s - s_gpu / J).max() <= 2e-6 * scale
    # the implicit Newton solution agrees to its own convergence
    _, _, s_impl, _ = _golden_sigma(golden, "implicit_last")
    assert np.abs(s - s_impl / J).max() <= 2e-6 * scale


def test_network27_tangent_matches_golden(golden, net27):
    """Forward-difference tangent with h=1e-5 vs explicit_cpu_out.txt:33-38.
    The golden run differs in termination (load time + smooth step), so the
    two printed blocks differ by 3e-4 and explicit vs implicit by 7e-4."""
    F, J, _, C_gold = _golden_sigma(golden)
    o = Oracle([net27])
    o.run(F)
    Cm = o.computeMaterialStiffness()[0]
    err = np.abs(Cm - C_gold / J).max() / np.abs(C_gold / J).max()
    assert err < 5e-3
    _, _, _, C_impl = _golden_sigma(golden, "implicit_last")
    assert np.abs(Cm - C_impl / J).max() / np.abs(C_impl / J).max() < 5e-3
    # diagonal and the dominant normal couplings are close to minor-symmetric
    assert abs(Cm[0, 1] - Cm[1, 0]) / Cm[0, 0] < 2e-3


def test_network27_orientation_matches_summary(golden, net27):
    """Orientation tensor of the undeformed network (Network_2_7.txt.summary)."""
    o = Oracle([net27])
    om = o.orientation()[0]
    assert np.allclose(om, np.array(golden["network_2_7"]["summary_orientation"]),
                       atol=1.5e-6)
    assert math.isclose(np.trace(om), 1.0, rel_tol=1e-12)
    # 2D tensor for normal x: in-plane projections only, trace 1
    om2 = o.orientation(normal=[1.0, 0, 0])[0]
    assert abs(om2[0, 0]) < 1e-12 and math.isclose(np.trace(om2), 1.0, rel_tol=1e-12)


def test_fiber_force_closed_forms():
    lin = mf.Reaction(mf.MF_REACTION_LINEAR, 0.1, 2.0, 1.0)
    EA = 2.0 * 0.01 * math.pi
    assert fiber_force(lin, 1.0, 1.0) == 0.0
    assert math.isclose(fiber_force(lin, 1.0, 2.0), 3 * EA, rel_tol=1e-15)
    assert math.isclose(fiber_force(lin, 0.5, 1.0), 3 * EA, rel_tol=1e-15)
    nl = mf.Reaction(mf.MF_REACTION_NONLINEAR, 0.1, 2.0, 1.0, 1.2, 1.05, 1.0)
    alpha = EA / 1.2
    # toe region: alpha * lam * (exp(B*E) - 1)
    lam = 1.03
    E = 0.5 * (lam * lam - 1)
    assert math.isclose(fiber_force(nl, 1.0, lam), alpha * lam * math.expm1(1.2 * E),
                        rel_tol=1e-12)
    # continuity of the force and of its slope at both transitions
    for lt in (1.0, 1.05):
        eps = 1e-7
        fm, f0, fp = (fiber_force(nl, 1.0, lt + d) for d in (-eps, 0.0, eps))
        # no jump: the midpoint value matches f0 up to curvature terms
        assert abs(0.5 * (fp + fm) - f0) < 1e-4 * abs(fp - fm)
        sl = (fp - f0) / eps
        sr = (f0 - fm) / eps
        assert math.isclose(sl, sr, rel_tol=1e-4)
    # compression branch with lc = 1 is the linear law EA (lam - 1)
    assert math.isclose(fiber_force(nl, 1.0, 0.9), EA * (0.9 - 1.0), rel_tol=1e-12)


def test_nodal_mass_sum(net27):
    """sum_v m_v = rho*A*sum(L0) (row-sum lumping of linear truss elements)."""
    o = Oracle([net27])
    m = o.nodal_mass(0)
    X, fb = net27.coords, net27.fibers
    L0 = np.linalg.norm(X[fb[:, 1]] - X[fb[:, 0]], axis=1)
    r = net27.reactions[0]
    rhoA = r.density * math.pi * r.radius ** 2
    assert math.isclose(m.sum(), rhoA * L0.sum(), rel_tol=1e-13)
    exp = np.zeros(len(X))
    np.add.at(exp, fb[:, 0], 0.5 * rhoA * L0)
    np.add.at(exp, fb[:, 1], 0.5 * rhoA * L0)
    assert np.allclose(m, exp, rtol=1e-13)


def _fiber_virial(net, x):
    """sum_e (f_e/l_e) d_e (x) d_e with the linear law (equilibrium stress)."""
    fb = net.fibers
    d = x[fb[:, 1]] - x[fb[:, 0]]
    l = np.linalg.norm(d, axis=1)
    X = net.coords
    l0 = np.linalg.norm(X[fb[:, 1]] - X[fb[:, 0]], axis=1)
    r = net.reactions[0]
    EA = r.E * math.pi * r.radius ** 2
    lam = l / l0
    f = lam * EA * 0.5 * (lam ** 2 - 1)
    S = np.einsum("e,ei,ej->ij", f / l, d, d)
    return np.array(voigt(S))


def test_virial_identity_at_equilibrium(golden, net27):
    """At equilibrium the boundary virial equals the fiber virial."""
    F, J, _, _ = _golden_sigma(golden)
    o = Oracle([net27])
    s = o.run(F)[0]
    x, u, V = o.get_state()
    sv = _fiber_virial(net27, x) / V[0]
    assert np.abs(s - sv).max() < 1e-4 * np.abs(s).max()


def test_affine_patch_all_boundary():
    """Every node prescribed: the stress is the closed-form affine virial."""
    rng = np.random.default_rng(3)
    X = rng.uniform(-0.5, 0.5, (12, 3))
    fb = np.array([[i, j] for i in range(12) for j in range(i + 1, 12) if (i + j) % 3 == 0],
                  np.int32)
    net = mf.Network(X, fb, np.arange(12, dtype=np.int32),
                     [mf.Reaction(0, 0.01, 1e5, 1000.0)])
    F = np.array([[1.02, 0.01, 0.0], [0.0, 0.99, 0.005], [0.003, 0.0, 1.01]])
    o = Oracle([net])
    s = o.run(F)[0]
    x = X @ F.T
    sv = _fiber_virial(net, x) / np.linalg.det(F)
    assert np.allclose(s, sv, rtol=1e-12, atol=1e-12 * np.abs(sv).max())


def test_chain_closed_form():
    """A straight chain between the faces x=-0.5 and x=+0.5 under
    F = diag(lam, 1, 1): every fiber stretches by lam, so
    sigma_xx = lam*EA*(lam^2-1)/2 (V = lam) and the other entries vanish."""
    xs = np.array([-0.5, -0.31, -0.02, 0.2, 0.37, 0.5])
    X = np.stack([xs, np.zeros(6), np.zeros(6)], 1)
    fb = np.array([[i, i + 1] for i in range(5)], np.int32)
    r = mf.Reaction(0, 0.01, 1e6, 1000.0)
    net = mf.Network(X, fb, np.array([0, 5], np.int32), [r])
    lam = 1.04
    o = Oracle([net])
    s = o.run(np.diag([lam, 1.0, 1.0]))[0]
    EA = r.E * math.pi * r.radius ** 2
    assert math.isclose(s[0], lam * EA * (lam ** 2 - 1) / 2, rel_tol=1e-10)
    assert np.abs(s[1:]).max() < 1e-12 * abs(s[0])


def _perturbation(i, h):
    """Perturbation of Voigt direction i (BatchedFiberRVEAnalysisExplicit.h:
    338-390): a normal stretch D1 = (1 - 2h)^(-1/2), or a symmetric shear
    with principal stretches sqrt(l2), sqrt(l3), l2 = (1+h)/(1-h^2),
    l3 = (1-h)/(1-h^2)."""
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
def test_chain_tangent_closed_form(scheme):
    """Relax-mode tangent of the straight chain from the closed form, forward
    and central (the central scheme has no reference counterpart).  The
    relaxed chain is straight between its end nodes a and b, every fiber at
    the same stretch, so sigma = EA (|c|^2 - 1)/2 c c^T / V with c = x_b - x_a
    (reference length 1) and V = det F of the committed state (trial runs
    keep the volume).  Trials follow the reference sequence: a trial restores
    the displacement but keeps the coordinates it ended with
    (BatchedFiberRVEAnalysisExplicit.h:264-313), and the next trial applies
    its increment to those, so the end nodes move to x_c + (F_q - I) x_prev."""
    xs = np.array([-0.5, -0.31, -0.02, 0.2, 0.37, 0.5])
    X = np.stack([xs, np.zeros(6), np.zeros(6)], 1)
    fb = np.array([[i, i + 1] for i in range(5)], np.int32)
    r = mf.Reaction(0, 0.01, 1e6, 1000.0)
    net = mf.Network(X, fb, np.array([0, 5], np.int32), [r])
    EA = r.E * math.pi * r.radius ** 2
    F = np.array([[1.04, 0.02, 0.0], [0.01, 0.99, 0.0], [0.0, 0.0, 1.01]])
    V = np.linalg.det(F)

    def sigma(c):
        return np.array(voigt(EA * (c @ c - 1) / 2 * np.outer(c, c) / V))

    h = 1e-4
    central = scheme == mf.MF_TANGENT_CENTRAL
    o = Oracle([net], mf.default_params(tolerance=1e-12, tangent_scheme=scheme, fd_step=h))
    s = o.run(F)[0]
    xa, xb = F @ X[0], F @ X[5]
    s0 = sigma(xb - xa)
    assert np.abs(s - s0).max() <= 1e-12 * np.abs(s0).max()
    C = o.computeMaterialStiffness()[0]
    Cc, Cu = np.zeros((6, 6)), np.zeros((6, 6))
    pa, pb = xa, xb
    for i in range(6):
        chained, fresh = [], []
        for sg in ((1, -1) if central else (1,)):
            Fq = _perturbation(i, sg * h)
            pa, pb = xa + (Fq - np.eye(3)) @ pa, xb + (Fq - np.eye(3)) @ pb
            chained.append(sigma(pb - pa))
            fresh.append(sigma(Fq @ (xb - xa)))
        for Cx, t in ((Cc, chained), (Cu, fresh)):
            Cx[:, i] = (t[0] - t[1]) / (2 * h) if central else (t[0] - s0) / h
    scale = np.abs(Cc).max()
    assert np.abs(C - Cc).max() <= 1e-9 * scale
    if central:
        # the chaining is visible at O(h): the -h trial starts where the +h
        # trial ended
        assert np.abs(C - Cu).max() >= 0.2 * h * scale


def test_relax_converges_from_perturbed_state():
    """Relax mode drives a chain displaced off equilibrium back to it."""
    xs = np.array([-0.5, -0.25, 0.0, 0.25, 0.5])
    X = np.stack([xs, np.zeros(5), np.zeros(5)], 1)
    fb = np.array([[i, i + 1] for i in range(4)], np.int32)
    r = mf.Reaction(0, 0.01, 1e6, 1000.0)
    net = mf.Network(X, fb, np.array([0, 4], np.int32), [r])
    o = Oracle([net], mf.default_params(tolerance=1e-9))
    s0 = o.run(np.diag([1.02, 1.0, 1.0]))[0]
    x, _, _ = o.get_state()
    # interior nodes end equally spaced
    assert np.allclose(np.diff(x[:, 0]), 1.02 * 0.25, rtol=1e-8)
    assert o.last_steps[0] % 100 == 1  # exits right after a refresh step
    EA = r.E * math.pi * r.radius ** 2
    assert math.isclose(s0[0], 1.02 * EA * (1.02 ** 2 - 1) / 2, rel_tol=1e-7)


def test_ramp_mode_approaches_equilibrium(net27):
    """With enough ramp steps the ramp-mode stress approaches the relaxed one;
    halving the load halves the small-strain stress."""
    F = np.diag([1.002, 1.0, 0.999])
    o_relax = Oracle([net27], mf.default_params(tolerance=1e-9))
    s_relax = o_relax.run(F)[0]
    o_ramp = Oracle([net27], mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=20000))
    s_ramp = o_ramp.run(F)[0]
    assert np.abs(s_ramp - s_relax).max() < 2e-3 * np.abs(s_relax).max()
    Fh = np.diag([1.001, 1.0, 0.9995])
    o_half = Oracle([net27], mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=20000))
    s_half = o_half.run(Fh)[0]
    assert np.abs(2 * s_half - s_ramp).max() < 5e-3 * np.abs(s_ramp).max()


def test_trial_run_restores_displacement(net27):
    """update_coords=false restores u but keeps the trial coordinates (the
    reference's run(), BatchedFiberRVEAnalysisExplicit.h:264-268,309-318)."""
    F = np.diag([1.01, 1.0, 1.0])
    o = Oracle([net27])
    o.run(F)
    x0, u0, V0 = o.get_state()
    o.run(np.diag([1.001, 1.0, 1.0]), update_coords=False)
    x1, u1, V1 = o.get_state()
    assert np.array_equal(u0, u1) and np.array_equal(V0, V1)
    assert not np.array_equal(x0, x1)


def test_nan_on_degenerate_fiber():
    """A zero-length fiber gives NaN forces; the run reports it."""
    X = np.array([[-0.5, 0, 0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.5, 0, 0]])
    fb = np.array([[0, 1], [1, 2], [2, 3]], np.int32)
    net = mf.Network(X, fb, np.array([0, 3], np.int32), [mf.Reaction(0, 0.01, 1e6, 1e3)])
    o = Oracle([net])
    s, st = o.run(np.diag([1.01, 1, 1]), return_status=True)
    assert st[0] == mf.MF_STATUS_NAN


# ---------------------------------------------------------------- diagnostics
def test_smoothstep_amplitude_known_values():
    """SmoothAmp (ExplicitAmplitude.h:26-40): xi^3 (10 - 15 xi + 6 xi^2)."""
    from oracle.oracle import ramp_alpha
    S = 1000
    assert ramp_alpha(0, S, mf.MF_AMP_SMOOTHSTEP) == 0.0
    assert ramp_alpha(S // 2, S, mf.MF_AMP_SMOOTHSTEP) == 0.5
    assert ramp_alpha(S, S, mf.MF_AMP_SMOOTHSTEP) == 1.0
    assert ramp_alpha(1, S, mf.MF_AMP_SMOOTHSTEP) == pytest.approx(10 / S ** 3, rel=2e-3)
    for k in (1, 37, 250, 499):  # point symmetry about xi = 1/2
        s = ramp_alpha(k, S, mf.MF_AMP_SMOOTHSTEP) + ramp_alpha(S - k, S, mf.MF_AMP_SMOOTHSTEP)
        assert abs(s - 1.0) < 1e-14
        assert ramp_alpha(k, S, mf.MF_AMP_LINEAR) == k / S


def test_smoothstep_ramp_ends_at_the_linear_ramp_state(net27):
    F = np.diag([1.02, 0.99, 1.0])
    out = []
    for amp in (mf.MF_AMP_LINEAR, mf.MF_AMP_SMOOTHSTEP):
        o = Oracle([net27], mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=20000,
                                               amplitude=amp))
        out.append(o.run(F)[0])
    assert np.abs(out[0] - out[1]).max() <= 2e-3 * np.abs(out[0]).max()


def _strain_energy(net, x):
    """sum over fibers of E A l0 (lambda^2 - 1)^2 / 8: the integral of the
    linear truss force lambda E A (lambda^2 - 1)/2 over the length."""
    r = net.reactions[0]
    EA = r.E * r.radius ** 2 * math.pi
    a, b = net.fibers[:, 0], net.fibers[:, 1]
    l0 = np.linalg.norm(net.coords[b] - net.coords[a], axis=1)
    l = np.linalg.norm(x[b] - x[a], axis=1)
    lam = l / l0
    return float(np.sum(EA * l0 * (lam ** 2 - 1) ** 2 / 8))


def test_energy_accounting_closed_forms(net27):
    """W_int tracks the strain energy of the final state, the balance
    W_kin + W_int + W_damp - W_ext closes, damping dissipates."""
    F = np.diag([1.03, 1.0, 0.99])
    p = mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=4000, energy_diagnostics=1)
    o = Oracle([net27], p)
    o.run(F)
    W_int, W_ext, W_damp, W_kin = o.last_energies()[0]
    x, _, _ = o.get_state()
    U = _strain_energy(net27, x)
    assert W_ext > 0 and W_damp > 0 and W_kin >= 0
    assert abs(W_int - U) <= 2e-3 * U, (W_int, U)
    assert abs(W_kin + W_int + W_damp - W_ext) <= 2e-3 * W_ext


def test_energy_check_status(net27):
    F = np.diag([1.03, 1.0, 1.0])
    for eps, want in ((1e-14, mf.MF_STATUS_ENERGY), (0.5, mf.MF_STATUS_OK)):
        p = mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=1000, energy_diagnostics=1,
                              energy_check_eps=eps)
        _, st = Oracle([net27], p).run(F, return_status=True)
        assert st[0] == want


def _affine_patch():
    """12 nodes, every one prescribed, linear fibers: the stress of any
    state the boundary reaches is the fiber virial of that affine state."""
    rng = np.random.default_rng(3)
    X = rng.uniform(-0.5, 0.5, (12, 3))
    fb = np.array([[i, j] for i in range(12) for j in range(i + 1, 12) if (i + j) % 3 == 0],
                  np.int32)
    return mf.Network(X, fb, np.arange(12, dtype=np.int32),
                      [mf.Reaction(0, 0.01, 1e5, 1000.0)])


@pytest.mark.parametrize("scheme", [mf.MF_TANGENT_FORWARD, mf.MF_TANGENT_CENTRAL])
def test_ramp_tangent_affine_patch_closed_form(scheme):
    """Ramp-mode tangent (no reference counterpart) on the all-boundary
    patch after two update runs F1 then F2.  A ramp solve moves every node
    to F x_start, so the perturbed solve i(+-h) ends at P_i(+-h) F2 F1 X and
    its stress is the closed-form virial of that state over the committed
    volume det F1 det F2.  The tangent must be the difference quotient of
    those closed forms: P_i applied after F2 (the perturbation is a further
    deformation of the current configuration,
    BatchedFiberRVEAnalysisExplicit.h:331-403), from the state before the
    last update run (F1 X).  Composing F2 P_i instead, or starting from X or
    from the committed state, moves C at O(|F - I|), which the test shows."""
    net = _affine_patch()
    X = net.coords
    F1 = np.array([[1.02, 0.01, 0.0], [0.0, 0.99, 0.005], [0.003, 0.0, 1.01]])
    F2 = np.array([[0.995, 0.0, 0.012], [0.004, 1.015, 0.0], [0.0, -0.008, 1.0]])
    h = 1e-6
    p = mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=200, tangent_scheme=scheme, fd_step=h)
    o = Oracle([net], p)
    o.run(F1)
    s_c = o.run(F2)[0]
    V = np.linalg.det(F1) * np.linalg.det(F2)

    def sig(x):
        return _fiber_virial(net, x) / V

    x1 = X @ F1.T
    s0 = sig(x1 @ F2.T)
    assert np.allclose(s_c, s0, rtol=0, atol=1e-12 * np.abs(s0).max())
    C = o.computeMaterialStiffness()[0]
    central = scheme == mf.MF_TANGENT_CENTRAL

    def closed(order, start):
        Cx = np.zeros((6, 6))
        for i in range(6):
            t = []
            for sg in ((1, -1) if central else (1,)):
                Pq = _perturbation(i, sg * h)
                Ft = Pq @ F2 if order == "PF" else F2 @ Pq
                t.append(sig(start @ Ft.T))
            Cx[:, i] = (t[0] - t[1]) / (2 * h) if central else (t[0] - s0) / h
        return Cx

    Cc = closed("PF", x1)
    scale = np.abs(Cc).max()
    assert np.abs(C - Cc).max() <= 1e-7 * scale, np.abs(C - Cc).max() / scale
    # the plausible mistakes are visible
    for wrong in (closed("FP", x1), closed("PF", X), closed("PF", x1 @ F2.T)):
        assert np.abs(C - wrong).max() >= 1e-3 * scale


def test_orientation_2d_non_unit_normal_known_answer():
    """compute2DOrientationTensor divides the normal by |n|^2, not |n|
    (BatchedComputeOrientationTensor.h:76-84), so a non-unit normal is not a
    projection.  Two fibers, along (1,0,1)/sqrt2 and (0,1,0), worked by hand:
    n = (0,0,2) scales to (0,0,1/2) and moves the first fiber's direction to
    (1,0,1-1/4)/|.| = (0.8,0,0.6); n = (0,0,1/2) scales to (0,0,2) and moves
    it to (1,0,1-4)/|.| = (1,0,-3)/sqrt10.  The second fiber is orthogonal to
    n either way."""
    X = np.array([[-0.25, 0.0, -0.25], [0.25, 0.0, 0.25], [0.1, -0.3, 0.0], [0.1, 0.3, 0.0]])
    fb = np.array([[0, 1], [2, 3]], np.int32)
    net = mf.Network(X, fb, np.array([0, 1, 2, 3], np.int32), [mf.Reaction(0, 0.01, 1e5, 1e3)])
    o = Oracle([net])
    e2 = np.diag([0.0, 1.0, 0.0])
    want = {2.0: (np.array([[0.64, 0, 0.48], [0, 0, 0], [0.48, 0, 0.36]]) + e2) / 2,
            0.5: (np.array([[0.1, 0, -0.3], [0, 0, 0], [-0.3, 0, 0.9]]) + e2) / 2,
            1.0: (np.diag([1.0, 0.0, 0.0]) + e2) / 2}
    for c, w in want.items():
        om = o.orientation(normal=[0.0, 0.0, c])[0]
        assert np.allclose(om, w, rtol=0, atol=1e-14), (c, om)


def test_tangent_linear_affine_patch_closed_form():
    """Oracle tangent-linear (ramp mode) on the all-boundary patch after two
    update runs: the derivative of the closed-form affine stress along
    P_i(h) F2 at h = 0, taken here by a central difference of the closed
    form at h = 1e-5 (its truncation, O(h^2), is ~1e-9 relative: it falls
    100-fold from h = 1e-4)."""
    net = _affine_patch()
    X = net.coords
    F1 = np.array([[1.02, 0.01, 0.0], [0.0, 0.99, 0.005], [0.003, 0.0, 1.01]])
    F2 = np.array([[0.995, 0.0, 0.012], [0.004, 1.015, 0.0], [0.0, -0.008, 1.0]])
    p = mf.default_params(mode=mf.MF_MODE_RAMP, n_steps=200, tangent_scheme=mf.MF_TANGENT_LINEAR)
    o = Oracle([net], p)
    o.run(F1)
    o.run(F2)
    C = o.computeMaterialStiffness()[0]
    V = np.linalg.det(F1) * np.linalg.det(F2)
    x1 = X @ F1.T
    h = 1e-5
    Cc = np.zeros((6, 6))
    for i in range(6):
        sp = _fiber_virial(net, x1 @ (_perturbation(i, h) @ F2).T) / V
        sm = _fiber_virial(net, x1 @ (_perturbation(i, -h) @ F2).T) / V
        Cc[:, i] = (sp - sm) / (2 * h)
    scale = np.abs(Cc).max()
    assert np.abs(C - Cc).max() <= 1e-8 * scale, np.abs(C - Cc).max() / scale


@pytest.mark.parametrize("nl", [0.0, 0.5])
def test_tangent_linear_matches_central_difference(nl):
    """On a generated 300-node network (linear, or half the fibers on the toe
    law) the oracle's tangent-linear equals its own central difference of
    whole ramp solves at h = 1e-6 (FP64) to the difference's accuracy:
    every term of the solve (ramp, kicks, forces, the refreshed critical
    step, the boundary sum) is differentiated."""
    net = mf.generate_network(seed=41, n_nodes=300, E=1e8, radius=5e-5, nonlinear_fraction=nl)
    F = np.array([[1.03, 0.01, -0.004], [0.01, 0.985, 0.002], [-0.004, 0.002, 1.012]])
    kw = dict(mode=mf.MF_MODE_RAMP, n_steps=1500, reaction_policy=mf.MF_REACT_PER_FIBER)
    ot = Oracle([net], mf.default_params(tangent_scheme=mf.MF_TANGENT_LINEAR, **kw))
    ot.run(F)
    Ct = ot.computeMaterialStiffness()[0]
    oc = Oracle([net], mf.default_params(tangent_scheme=mf.MF_TANGENT_CENTRAL, fd_step=1e-6, **kw))
    oc.run(F)
    Cc = oc.computeMaterialStiffness()[0]
    err = np.abs(Ct - Cc).max() / np.abs(Cc).max()
    assert err <= 1e-7, err
