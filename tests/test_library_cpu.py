"""Host-side checks of libmumfim_amd.so that need no GPU: the C ABI exports,
the seeded generator and the reference-format loader."""
import os
import re

import numpy as np
import pytest

import mumfim_amd as mf
from conftest import ROOT, GOLDEN


def header_functions():
    txt = open(os.path.join(ROOT, "include", "mumfim_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mf_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = mf.lib()
    names = header_functions()
    assert len(names) > 30
    for n in names:
        assert hasattr(lib, n), n
    # the Python signature table covers the whole header
    assert sorted(mf.SIGNATURES) == names
    assert lib.mf_version() >= 1


def test_default_params_reproduce_reference_constants():
    p = mf.default_params()
    assert p.mode == mf.MF_MODE_RELAX and p.precision == mf.MF_PREC_FP64
    assert p.dt_update_freq == 100  # UPDATE_FREQ, TeamOuterLoop.h:4
    assert p.crit_time_scale_factor == 0.8  # TeamOuterLoop.h:312
    assert p.tolerance == 1e-6  # TeamOuterLoop.h:371
    assert p.fd_step == 1e-5  # BatchedFiberRVEAnalysisExplicit.h:331
    assert p.visc_damp_coeff == 100.0  # fiber_only.yaml:16


def _knn_bruteforce(X, k):
    d = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d, np.inf)
    edges = set()
    for i in range(len(X)):
        order = np.lexsort((np.arange(len(X)), d[i]))[:k]
        for j in order:
            edges.add((min(i, j), max(i, j)))
    return np.array(sorted(edges), np.int32)


@pytest.mark.parametrize("n,seed", [(60, 1), (200, 7), (500, 3)])
def test_generator_is_exact_knn(n, seed):
    net = mf.generate_network(seed=seed, n_nodes=n)
    X = net.coords
    assert X.shape == (n, 3) and X.min() >= -0.5 and X.max() < 0.5
    ref = _knn_bruteforce(X, 4)
    assert np.array_equal(net.fibers, ref)
    band = 0.5 - np.abs(X).max(1) <= 0.02
    assert np.array_equal(net.boundary, np.nonzero(band)[0])
    deg = np.bincount(net.fibers.ravel(), minlength=n)
    assert deg.min() >= 4
    # ~2.5 fibers per node for the symmetric 4-NN graph
    assert 2.0 < len(net.fibers) / n < 3.0


def test_generator_is_deterministic_and_seeded():
    a = mf.generate_network(seed=11, n_nodes=300, nonlinear_fraction=0.5)
    b = mf.generate_network(seed=11, n_nodes=300, nonlinear_fraction=0.5)
    c = mf.generate_network(seed=12, n_nodes=300, nonlinear_fraction=0.5)
    assert np.array_equal(a.coords, b.coords) and np.array_equal(a.fibers, b.fibers)
    assert np.array_equal(a.fiber_reaction, b.fiber_reaction)
    assert not np.array_equal(a.coords, c.coords)
    frac = a.fiber_reaction.mean()
    assert 0.4 < frac < 0.6
    assert a.reactions[1].type == mf.MF_REACTION_NONLINEAR
    assert a.reactions[1].B == 1.2 and a.reactions[1].length_ratio_trns == 1.05


def test_uniform_stream():
    u = mf.uniform(5, 9, 1000)
    assert np.array_equal(u, mf.uniform(5, 9, 1000))
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.05
    assert not np.array_equal(u, mf.uniform(5, 10, 1000))


def test_load_reference_network():
    net = mf.load_network(os.path.join(GOLDEN, "network_2_7", "Network_2_7.txt"))
    assert (net.n_nodes, net.n_fibers) == (336, 635)  # Network_2_7.txt.summary
    assert len(net.boundary) == 261
    r = net.reactions[0]
    assert (r.type, r.radius, r.E, r.density) == (0, 5e-5, 1e8, 1000.0)
    assert np.all(net.fiber_reaction == 0)
    on = (np.abs(np.abs(net.coords) - 0.5) < 1e-8).any(1)
    assert np.array_equal(net.boundary, np.nonzero(on)[0])


def test_load_errors_are_reported():
    with pytest.raises(mf.MfError):
        mf.load_network(os.path.join(GOLDEN, "does_not_exist.txt"))


def test_batch_create_rejects_bad_input_before_touching_the_gpu():
    net = mf.generate_network(seed=1, n_nodes=50)
    bad = mf.Network(net.coords, np.array([[0, 0]], np.int32), net.boundary, net.reactions)
    with pytest.raises(mf.MfError):
        mf.BatchedFiberRVEAnalysisExplicit([bad])


def test_planner_takes_any_rve_size():
    """An RVE with more free nodes than any register/LDS kernel holds (12000
    nodes) is planned onto the FP64 global-state kernel in both precisions,
    so creation gets past planning (here it stops at the missing device);
    with FP64 energy diagnostics, which that kernel does not track, it is
    refused by the planner."""
    net = mf.generate_network(seed=3, n_nodes=12000, nonlinear_fraction=0.5)
    for prec in (mf.MF_PREC_FP64, mf.MF_PREC_FP32C):
        try:
            mf.BatchedFiberRVEAnalysisExplicit([net], mf.default_params(precision=prec)).close()
        except mf.MfError as e:
            assert "fits no kernel variant" not in str(e), e
    with pytest.raises(mf.MfError, match="fits no kernel variant"):
        mf.BatchedFiberRVEAnalysisExplicit([net], mf.default_params(energy_diagnostics=1))


def test_load_reference_network_del_4450():
    """The reference's largest test network (12308 nodes, 38376 fibers, one
    reaction from its .params file) spans [-2.5, 2.5]^3: no node on the
    [-0.5, 0.5]^3 RVE faces, so the loader finds no boundary node."""
    net = mf.load_network(os.path.join(GOLDEN, "del_4450", "del_4450seedL5_new_1.txt"))
    assert (net.n_nodes, net.n_fibers, len(net.boundary)) == (12308, 38376, 0)
    assert np.isclose(np.abs(net.coords).max(), 2.5)
    assert len(net.reactions) == 1 and net.reactions[0].E == 1e8


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus above 1 exits with status 2 and says
    so, before it loads the library or touches a device."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_header_is_plain_c_and_links(tmp_path):
    """include/mumfim_amd.h compiles as C99 with -Wall -Werror (no C++
    types cross the boundary), and a C program that calls entry points
    links against libmumfim_amd.so. Nothing runs: no device is touched."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    src = tmp_path / "use.c"
    src.write_text(
        '#include "mumfim_amd.h"\n'
        "int main(void) {\n"
        "  mf_solver_params p;\n"
        "  mf_default_params(&p);\n"
        "  mf_batch *b = 0;\n"
        "  int rc = mf_batch_create(0, 0, &p, &b);\n"
        "  return rc == MF_OK ? 1 : (mf_last_error() != 0 ? 0 : 2);\n"
        "}\n")
    lib = os.path.join(ROOT, "mumfim_amd")
    r = subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        str(src), "-o", str(tmp_path / "use"), "-L", lib, "-lmumfim_amd",
                        f"-Wl,-rpath,{lib}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_wrappers_copy_the_params_struct():
    """set_params on one batch never changes the struct the caller passed
    (and may pass to another batch or to the oracle)."""
    import mumfim_amd as mf
    from oracle.oracle import Oracle
    p = mf.default_params(fd_step=1e-6)
    net = mf.generate_network(seed=3, n_nodes=60)
    o = Oracle([net], p)
    o.set_params(fd_step=1e-5)
    assert p.fd_step == 1e-6 and o.params.fd_step == 1e-5
