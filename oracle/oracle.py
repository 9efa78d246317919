"""TEST INFRASTRUCTURE ONLY: ctypes view of the CPU FP64 oracle
(oracle/mf_oracle.cpp).  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg; never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from mumfim_amd import (mf_network, mf_reaction, mf_solver_params, _dptr, _iptr,
                        default_params)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = C.CDLL(LIB_PATH)
        P = C.c_void_p
        D = C.POINTER(C.c_double)
        I = C.POINTER(C.c_int32)
        l.oracle_create.argtypes = [C.POINTER(mf_network), C.c_int,
                                    C.POINTER(mf_solver_params), C.POINTER(P)]
        l.oracle_destroy.argtypes = [P]
        l.oracle_set_params.argtypes = [P, C.POINTER(mf_solver_params)]
        l.oracle_run.argtypes = [P, D, C.c_int, D, C.c_int, I, I]
        l.oracle_tangent.argtypes = [P, D]
        l.oracle_get_state.argtypes = [P, D, D, D]
        l.oracle_orientation.argtypes = [P, D, C.c_int, D]
        l.oracle_reorder_permutation.argtypes = [C.c_int, I, C.c_int, I]
        l.oracle_fiber_force.argtypes = [C.POINTER(mf_reaction), C.c_double, C.c_double]
        l.oracle_fiber_force.restype = C.c_double
        l.oracle_nodal_mass.argtypes = [P, C.c_int, D]
        l.oracle_last_energies.argtypes = [P, D]
        l.oracle_ramp_alpha.argtypes = [C.c_int, C.c_int, C.c_int]
        l.oracle_ramp_alpha.restype = C.c_double
        _lib = l
    return _lib


class Oracle:
    """Mirror of BatchedFiberRVEAnalysisExplicit on the CPU oracle."""

    def __init__(self, networks, params=None):
        self.params = (mf_solver_params.from_buffer_copy(params) if params is not None
                       else default_params())
        self.nets = networks
        arr = (mf_network * len(networks))(*[n.to_c() for n in networks])
        h = C.c_void_p()
        lib().oracle_create(arr, len(networks), C.byref(self.params), C.byref(h))
        self.h = h
        self.n = len(networks)
        self.n_nodes = sum(n.n_nodes for n in networks)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        lib().oracle_set_params(self.h, C.byref(self.params))

    def run(self, F, update_coords=True, return_status=False):
        F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(-1, 9))
        sigma = np.zeros((self.n, 6))
        st = np.zeros(self.n, np.int32)
        steps = np.zeros(self.n, np.int32)
        lib().oracle_run(self.h, _dptr(F), F.shape[0], _dptr(sigma),
                         int(bool(update_coords)), _iptr(st), _iptr(steps))
        self.last_steps = steps
        return (sigma, st) if return_status else sigma

    def last_energies(self):
        W = np.zeros((self.n, 4))
        lib().oracle_last_energies(self.h, W.ctypes.data_as(C.POINTER(C.c_double)))
        return W

    def computeMaterialStiffness(self):
        Cm = np.zeros((self.n, 6, 6))
        lib().oracle_tangent(self.h, _dptr(Cm))
        return Cm

    def get_state(self):
        x = np.zeros((self.n_nodes, 3))
        u = np.zeros((self.n_nodes, 3))
        V = np.zeros(self.n)
        lib().oracle_get_state(self.h, _dptr(x), _dptr(u), _dptr(V))
        return x, u, V

    def orientation(self, normal=None):
        om = np.zeros((self.n, 3, 3))
        if normal is None:
            lib().oracle_orientation(self.h, None, 0, _dptr(om))
        else:
            nv = np.ascontiguousarray(np.asarray(normal, dtype=np.float64).reshape(-1, 3))
            lib().oracle_orientation(self.h, _dptr(nv), nv.shape[0], _dptr(om))
        return om

    def nodal_mass(self, rve):
        m = np.zeros(self.nets[rve].n_nodes)
        lib().oracle_nodal_mass(self.h, rve, _dptr(m))
        return m


def ramp_alpha(k, S, amplitude):
    return lib().oracle_ramp_alpha(k, S, amplitude)


def reorder_permutation(nvert, fixed):
    fixed = np.ascontiguousarray(fixed, dtype=np.int32)
    out = np.zeros(nvert, np.int32)
    lib().oracle_reorder_permutation(nvert, _iptr(fixed), len(fixed), _iptr(out))
    return out


def fiber_force(reaction, l0, l):
    r = reaction.to_c()
    return lib().oracle_fiber_force(C.byref(r), l0, l)
