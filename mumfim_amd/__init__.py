"""mumfim_amd: batched explicit fiber-network RVE solver on MI355X (gfx950).

Python view of the C ABI in include/mumfim_amd.h (ctypes).  The compute
path is the HIP library libmumfim_amd.so; this module only marshals arrays.
`BatchedFiberRVEAnalysisExplicit` mirrors the reference's batched analysis
(src/mumfim/microscale/BatchedFiberRVEAnalysisExplicit.h:28-417):
run(F, sigma, update_coords), computeMaterialStiffness(C),
compute3DOrientationTensor(omega), compute2DOrientationTensor(normal, omega).
"""
import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmumfim_amd.so")

MF_OK = 0
MF_MODE_RELAX, MF_MODE_RAMP = 0, 1
MF_PREC_FP64, MF_PREC_FP32C = 0, 1
MF_REACT_BATCHED_REFERENCE, MF_REACT_PER_FIBER = 0, 1
MF_TANGENT_FORWARD, MF_TANGENT_CENTRAL, MF_TANGENT_LINEAR = 0, 1, 2
MF_REACTION_LINEAR, MF_REACTION_NONLINEAR = 0, 1
MF_STATUS_OK, MF_STATUS_NAN, MF_STATUS_STEP_CAP, MF_STATUS_FALLBACK = 0, 1, 2, 3
MF_STATUS_ENERGY = 4
MF_AMP_LINEAR, MF_AMP_SMOOTHSTEP = 0, 1
MF_FP32_KERNEL_AUTO, MF_FP32_KERNEL_NODE, MF_FP32_KERNEL_FIBER, MF_FP32_KERNEL_WAVE = 0, 1, 2, 3


class mf_reaction(C.Structure):
    _fields_ = [("type", C.c_int32), ("pad_", C.c_int32), ("radius", C.c_double),
                ("E", C.c_double), ("density", C.c_double), ("B", C.c_double),
                ("length_ratio_trns", C.c_double), ("cmp_ratio_trns", C.c_double)]


class mf_network(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("n_fibers", C.c_int32),
                ("n_boundary", C.c_int32), ("n_reactions", C.c_int32),
                ("coords", C.POINTER(C.c_double)), ("fibers", C.POINTER(C.c_int32)),
                ("fiber_reaction", C.POINTER(C.c_int32)),
                ("boundary", C.POINTER(C.c_int32)),
                ("reactions", C.POINTER(mf_reaction)), ("scale_factor", C.c_double)]


class mf_solver_params(C.Structure):
    _fields_ = [("mode", C.c_int32), ("precision", C.c_int32),
                ("reaction_policy", C.c_int32), ("tangent_scheme", C.c_int32),
                ("tangent_precision", C.c_int32), ("n_steps", C.c_int32),
                ("max_steps", C.c_int32), ("dt_update_freq", C.c_int32),
                ("block_size", C.c_int32), ("amplitude", C.c_int32),
                ("visc_damp_coeff", C.c_double),
                ("crit_time_scale_factor", C.c_double),
                ("tolerance", C.c_double), ("fd_step", C.c_double),
                ("energy_diagnostics", C.c_int32), ("fp32_kernel", C.c_int32),
                ("energy_check_eps", C.c_double), ("chunks", C.c_int32), ("pad_", C.c_int32)]


class mf_gen_params(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_nodes", C.c_int32),
                ("k_neighbors", C.c_int32), ("boundary_band", C.c_double),
                ("radius", C.c_double), ("E", C.c_double), ("density", C.c_double),
                ("nonlinear_fraction", C.c_double), ("B", C.c_double),
                ("length_ratio_trns", C.c_double)]


MF_MLP_IN, MF_MLP_HIDDEN, MF_MLP_OUT, MF_MLP_MAX_MEMBERS, MF_MLP_BATCH_ALIGN = 10, 256, 6, 8, 256


class mf_train_params(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("seed", C.c_uint64),
                ("lr", C.c_float), ("lr_final", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float)]


class mf_estimator(C.Structure):
    _fields_ = [("in_lo", C.c_float * MF_MLP_IN), ("in_hi", C.c_float * MF_MLP_IN),
                ("margin", C.c_float), ("kappa", C.c_float), ("sigma_floor", C.c_float),
                ("tol", C.c_float)]


class mf_surrogate_stats(C.Structure):
    _fields_ = [("n", C.c_int64), ("n_fallback", C.c_int64), ("n_out_of_range", C.c_int64),
                ("mlp_ms", C.c_double), ("fallback_ms", C.c_double), ("total_ms", C.c_double)]


# Every exported symbol of include/mumfim_amd.h with its ctypes signature.
_P = C.c_void_p
_I = C.c_int32
_D = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_FP = C.POINTER(C.c_float)
SIGNATURES = {
    "mf_last_error": (C.c_char_p, []),
    "mf_version": (C.c_int, []),
    "mf_default_params": (None, [C.POINTER(mf_solver_params)]),
    "mf_batch_create": (C.c_int, [C.POINTER(mf_network), _I, C.POINTER(mf_solver_params), C.POINTER(_P)]),
    "mf_batch_destroy": (None, [_P]),
    "mf_batch_num_rves": (C.c_int, [_P]),
    "mf_batch_set_params": (C.c_int, [_P, C.POINTER(mf_solver_params)]),
    "mf_batch_get_params": (C.c_int, [_P, C.POINTER(mf_solver_params)]),
    "mf_batch_run": (C.c_int, [_P, _D, _I, _D, _I, _IP]),
    "mf_batch_tangent": (C.c_int, [_P, _D, _IP]),
    "mf_batch_run_with_tangent": (C.c_int, [_P, _D, _I, _D, _D, _IP]),
    "mf_batch_run_device_async": (C.c_int, [_P, _P, _I, _P, _I, _P]),
    "mf_batch_run_device_noalloc": (C.c_int, [_P, _P, _I, _P, _I, _P]),
    "mf_batch_synchronize": (C.c_int, [_P]),
    "mf_batch_stream": (C.c_void_p, [_P]),
    "mf_batch_last_solve_ms": (C.c_int, [_P, _D, _IP]),
    "mf_batch_last_energies": (C.c_int, [_P, _D]),
    "mf_batch_autotune": (C.c_int, [_P, _I, _I, C.c_char_p, _IP]),
    "mf_batch_set_graph": (C.c_int, [_P, _I]),
    "mf_batch_orientation_3d": (C.c_int, [_P, _D]),
    "mf_batch_orientation_2d": (C.c_int, [_P, _D, _I, _D]),
    "mf_batch_total_nodes": (C.c_int64, [_P]),
    "mf_batch_get_state": (C.c_int, [_P, _D, _D, _D]),
    "mf_batch_set_state": (C.c_int, [_P, _D, _D, _D]),
    "mf_batch_plan": (C.c_int, [_P, _IP]),
    "mf_batch_rve_sizes": (C.c_int, [_P, _IP]),
    "mf_default_gen_params": (None, [C.POINTER(mf_gen_params)]),
    "mf_generate_network": (C.c_int, [C.POINTER(mf_gen_params), _IP, _IP, _IP, _D, _IP, _IP, _IP, C.POINTER(mf_reaction)]),
    "mf_batch_create_generated": (C.c_int, [C.POINTER(mf_gen_params), _I, _IP, C.POINTER(mf_solver_params), C.POINTER(_P)]),
    "mf_uniform": (C.c_int, [C.c_uint64, C.c_uint64, C.c_int64, _D]),
    "mf_device_alloc": (C.c_int, [C.POINTER(_P), C.c_int64]),
    "mf_device_free": (C.c_int, [_P]),
    "mf_memcpy_to_device": (C.c_int, [_P, _P, C.c_int64]),
    "mf_memcpy_to_host": (C.c_int, [_P, _P, C.c_int64]),
    "mf_device_synchronize": (C.c_int, []),
    "mf_measure_lds_bandwidth": (C.c_int, [_D]),
    "mf_measure_hbm_bandwidth": (C.c_int, [C.c_int64, _D]),
    "mf_load_network": (C.c_int, [C.c_char_p, _IP, _IP, _IP, _IP, _I, _D, _IP, _IP, _IP, C.POINTER(mf_reaction)]),
    "mf_mlp_create": (C.c_int, [_I, C.c_uint64, C.POINTER(_P)]),
    "mf_mlp_destroy": (None, [_P]),
    "mf_mlp_num_params": (C.c_int32, []),
    "mf_mlp_num_members": (C.c_int32, [_P]),
    "mf_mlp_set_normalization": (C.c_int, [_P, _FP, _FP, _FP, _FP]),
    "mf_mlp_get_normalization": (C.c_int, [_P, _FP, _FP, _FP, _FP]),
    "mf_mlp_get_weights": (C.c_int, [_P, _I, _FP]),
    "mf_mlp_set_weights": (C.c_int, [_P, _I, _FP]),
    "mf_mlp_train_step_device": (C.c_int, [_P, _P, _P, C.c_int64, _P, _I, C.c_float, _FP]),
    "mf_default_train_params": (None, [C.POINTER(mf_train_params)]),
    "mf_mlp_fit_device": (C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(mf_train_params), _FP]),
    "mf_mlp_gradient_device": (C.c_int, [_P, _I, _P, _P, _I, _FP, _FP]),
    "mf_mlp_predict_device": (C.c_int, [_P, _I, _P, C.c_int64, _P, _P, _P]),
    "mf_mlp_last_predict_ms": (C.c_int, [_P, _D]),
    "mf_mlp_stream": (C.c_void_p, [_P]),
    "mf_mlp_set_estimator": (C.c_int, [_P, C.POINTER(mf_estimator)]),
    "mf_mlp_get_estimator": (C.c_int, [_P, C.POINTER(mf_estimator)]),
    "mf_surrogate_inputs": (C.c_int, [_P, _D, _I, _FP]),
    "mf_batch_set_damping": (C.c_int, [_P, _D]),
    "mf_fo_scale_conversion": (C.c_int, [C.POINTER(mf_network), C.c_double, C.c_double, _D]),
    "mf_fo_create": (C.c_int, [C.POINTER(mf_network), _I, _IP, _D, _D, _IP,
                               C.POINTER(mf_solver_params), C.POINTER(_P)]),
    "mf_fo_destroy": (None, [_P]),
    "mf_fo_batch": (C.c_void_p, [_P]),
    "mf_fo_run": (C.c_int, [_P, _D, _D, _IP]),
    "mf_fo_step_results": (C.c_int, [_P, _D]),
    "mf_surrogate_run": (C.c_int, [_P, _P, _D, _I, _D, _D, _IP, C.POINTER(mf_surrogate_stats)]),
}

_lib = None


def lib():
    """Loads libmumfim_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m mumfim_amd.build` "
                "(the solver has no CPU fallback)")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


class MfError(RuntimeError):
    pass


def check(rc):
    if rc != MF_OK:
        raise MfError(f"mumfim_amd error {rc}: {lib().mf_last_error().decode()}")
    return rc


def _dptr(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _iptr(a):
    return a.ctypes.data_as(_IP) if a is not None else None


@dataclass
class Reaction:
    type: int = MF_REACTION_LINEAR
    radius: float = 0.0
    E: float = 0.0
    density: float = 0.0
    B: float = 0.0
    length_ratio_trns: float = 0.0
    cmp_ratio_trns: float = 1.0

    def to_c(self):
        return mf_reaction(self.type, 0, self.radius, self.E, self.density, self.B,
                           self.length_ratio_trns, self.cmp_ratio_trns)


@dataclass
class Network:
    """One fiber network (one RVE) in input node order."""
    coords: np.ndarray                 # (N, 3) float64
    fibers: np.ndarray                 # (M, 2) int32
    boundary: np.ndarray               # (Nb,) int32
    reactions: List[Reaction]
    fiber_reaction: Optional[np.ndarray] = None  # (M,) int32
    scale_factor: float = 1.0
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_nodes(self):
        return int(self.coords.shape[0])

    @property
    def n_fibers(self):
        return int(self.fibers.shape[0])

    def to_c(self):
        self.coords = np.ascontiguousarray(self.coords, dtype=np.float64)
        self.fibers = np.ascontiguousarray(self.fibers, dtype=np.int32)
        self.boundary = np.ascontiguousarray(self.boundary, dtype=np.int32)
        fr = None
        if self.fiber_reaction is not None:
            self.fiber_reaction = np.ascontiguousarray(self.fiber_reaction, dtype=np.int32)
            fr = self.fiber_reaction
        rx = (mf_reaction * len(self.reactions))(*[r.to_c() for r in self.reactions])
        self._keep = [rx]
        return mf_network(self.n_nodes, self.n_fibers, len(self.boundary),
                          len(self.reactions), _dptr(self.coords), _iptr(self.fibers),
                          _iptr(fr), _iptr(self.boundary), rx, self.scale_factor)


def _reaction_from_c(r):
    return Reaction(r.type, r.radius, r.E, r.density, r.B, r.length_ratio_trns,
                    r.cmp_ratio_trns)


def default_gen_params(**kw):
    g = mf_gen_params()
    lib().mf_default_gen_params(C.byref(g))
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def generate_network(seed=1, n_nodes=500, **kw) -> Network:
    """Seeded synthetic network: uniform nodes, deduplicated k-NN fibers,
    boundary band 0.02 (see mf_generate_network)."""
    g = default_gen_params(seed=seed, n_nodes=n_nodes, **kw)
    n, m, nb = C.c_int32(), C.c_int32(), C.c_int32()
    check(lib().mf_generate_network(C.byref(g), C.byref(n), C.byref(m), C.byref(nb),
                                    None, None, None, None, None))
    X = np.zeros((n.value, 3))
    fb = np.zeros((m.value, 2), np.int32)
    fr = np.zeros(m.value, np.int32)
    bd = np.zeros(nb.value, np.int32)
    rx = (mf_reaction * 2)()
    check(lib().mf_generate_network(C.byref(g), C.byref(n), C.byref(m), C.byref(nb),
                                    _dptr(X), _iptr(fb), _iptr(fr), _iptr(bd), rx))
    return Network(X, fb, bd, [_reaction_from_c(r) for r in rx], fr)


def load_network(path) -> Network:
    """Reads the reference's network text format and its .params file."""
    n, m, nb, nr = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
    p = str(path).encode()
    check(lib().mf_load_network(p, C.byref(n), C.byref(m), C.byref(nb), C.byref(nr), 0,
                                None, None, None, None, None))
    X = np.zeros((n.value, 3))
    fb = np.zeros((m.value, 2), np.int32)
    fr = np.zeros(m.value, np.int32)
    bd = np.zeros(nb.value, np.int32)
    rx = (mf_reaction * nr.value)()
    check(lib().mf_load_network(p, C.byref(n), C.byref(m), C.byref(nb), C.byref(nr),
                                nr.value, _dptr(X), _iptr(fb), _iptr(fr), _iptr(bd), rx))
    return Network(X, fb, bd, [_reaction_from_c(r) for r in rx], fr)


def uniform(seed, stream, n):
    out = np.zeros(n)
    check(lib().mf_uniform(seed, stream, n, _dptr(out)))
    return out


class DeviceArray:
    """An HBM buffer owned by the library's HIP runtime (float64, float32 or
    int32), with host copies in and out."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().mf_device_alloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.shape, a.dtype)
        d.copy_from(a)
        return d

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().mf_memcpy_to_device(self.ptr, a.ctypes.data, self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().mf_memcpy_to_host(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def numel(self):
        return int(np.prod(self.shape))

    def data_ptr(self):
        return self.ptr

    def free(self):
        if self.ptr:
            lib().mf_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_synchronize():
    check(lib().mf_device_synchronize())


def measure_lds_bandwidth():
    v = C.c_double()
    check(lib().mf_measure_lds_bandwidth(C.byref(v)))
    return v.value


def measure_hbm_bandwidth(nbytes=1 << 30):
    v = C.c_double()
    check(lib().mf_measure_hbm_bandwidth(nbytes, C.byref(v)))
    return v.value


def default_params(**kw):
    p = mf_solver_params()
    lib().mf_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class BatchedFiberRVEAnalysisExplicit:
    """Host mirror of the reference batched analysis over the C ABI."""

    def __init__(self, networks=None, params=None, *, generated=None, n_nodes=None):
        # own copy: set_params() must not change the caller's struct
        self.params = (mf_solver_params.from_buffer_copy(params) if params is not None
                       else default_params())
        h = C.c_void_p()
        if generated is not None:
            g, n_rve = generated
            nn = None
            if n_nodes is not None:
                self._nn = np.ascontiguousarray(n_nodes, dtype=np.int32)
                nn = _iptr(self._nn)
            check(lib().mf_batch_create_generated(C.byref(g), n_rve, nn,
                                                  C.byref(self.params), C.byref(h)))
        else:
            nets = (mf_network * len(networks))(*[n.to_c() for n in networks])
            check(lib().mf_batch_create(nets, len(networks), C.byref(self.params),
                                        C.byref(h)))
        self.h = h
        self.n = lib().mf_batch_num_rves(h)

    def close(self):
        if self.h:
            lib().mf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_damping(self, eta=None):
        """Per-RVE viscous damping (None: params.visc_damp_coeff for all)."""
        if eta is None:
            check(lib().mf_batch_set_damping(self.h, None))
        else:
            e = np.ascontiguousarray(eta, dtype=np.float64).reshape(self.n)
            check(lib().mf_batch_set_damping(self.h, _dptr(e)))

    def set_params(self, **kw):
        for k, v in kw.items():
            setattr(self.params, k, v)
        check(lib().mf_batch_set_params(self.h, C.byref(self.params)))

    def run(self, F, update_coords=True, return_status=False):
        F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(-1, 9))
        sigma = np.zeros((self.n, 6))
        status = np.zeros(self.n, np.int32)
        check(lib().mf_batch_run(self.h, _dptr(F), F.shape[0], _dptr(sigma),
                                 int(bool(update_coords)), _iptr(status)))
        return (sigma, status) if return_status else sigma

    def computeMaterialStiffness(self, return_status=False):
        Cm = np.zeros((self.n, 6, 6))
        status = np.zeros(self.n, np.int32)
        check(lib().mf_batch_tangent(self.h, _dptr(Cm), _iptr(status)))
        return (Cm, status) if return_status else Cm

    def run_with_tangent(self, F, return_status=False):
        F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(-1, 9))
        sigma = np.zeros((self.n, 6))
        Cm = np.zeros((self.n, 6, 6))
        status = np.zeros(self.n, np.int32)
        check(lib().mf_batch_run_with_tangent(self.h, _dptr(F), F.shape[0], _dptr(sigma),
                                              _dptr(Cm), _iptr(status)))
        return (sigma, Cm, status) if return_status else (sigma, Cm)

    def compute3DOrientationTensor(self):
        om = np.zeros((self.n, 3, 3))
        check(lib().mf_batch_orientation_3d(self.h, _dptr(om)))
        return om

    def compute2DOrientationTensor(self, normal):
        nv = np.ascontiguousarray(np.asarray(normal, dtype=np.float64).reshape(-1, 3))
        om = np.zeros((self.n, 3, 3))
        check(lib().mf_batch_orientation_2d(self.h, _dptr(nv), nv.shape[0], _dptr(om)))
        return om

    def get_state(self):
        nn = lib().mf_batch_total_nodes(self.h)
        x = np.zeros((nn, 3))
        u = np.zeros((nn, 3))
        V = np.zeros(self.n)
        check(lib().mf_batch_get_state(self.h, _dptr(x), _dptr(u), _dptr(V)))
        return x, u, V

    def autotune(self, precision=None, steps=200, cache_path=None):
        """Times the fitting kernel variants per RVE bucket and keeps the
        fastest (if the whole batch is then faster), then, for ramp-mode
        FP32c batches, the step chunks per solve; returns the number of
        timed kernel-variant candidates."""
        prec = self.params.precision if precision is None else precision
        n = C.c_int32()
        check(lib().mf_batch_autotune(self.h, int(prec), int(steps),
                                      None if cache_path is None else str(cache_path).encode(),
                                      C.byref(n)))
        return n.value

    def set_graph(self, enable=True):
        check(lib().mf_batch_set_graph(self.h, int(bool(enable))))

    def last_energies(self):
        """W_int, W_ext, W_damp, W_kin of the last run (energy diagnostics)."""
        W = np.zeros((self.n, 4))
        check(lib().mf_batch_last_energies(self.h, _dptr(W)))
        return W

    def last_solve_ms(self):
        ms = C.c_double()
        nl = C.c_int32()
        check(lib().mf_batch_last_solve_ms(self.h, C.byref(ms), C.byref(nl)))
        return ms.value, nl.value

    def stats(self):
        """Per-RVE node/fiber counts as numpy arrays N, Nf, Nb, M."""
        z = np.zeros((self.n, 4), np.int32)
        check(lib().mf_batch_rve_sizes(self.h, _iptr(z)))
        return {"N": z[:, 0].astype(np.int64), "Nf": z[:, 1].astype(np.int64),
                "Nb": z[:, 2].astype(np.int64), "M": z[:, 3].astype(np.int64)}

    def run_device(self, F_dev, sigma_dev, update_coords=False, status_dev=None):
        """Queues run() on device buffers (DeviceArray or raw pointers);
        F_dev holds 1 or n_rve row-major 3x3 matrices in float64."""
        fp = F_dev.data_ptr() if hasattr(F_dev, "data_ptr") else int(F_dev)
        sp = sigma_dev.data_ptr() if hasattr(sigma_dev, "data_ptr") else int(sigma_dev)
        nF = (F_dev.numel() // 9) if hasattr(F_dev, "numel") else self.n
        st = None
        if status_dev is not None:
            st = status_dev.data_ptr() if hasattr(status_dev, "data_ptr") else int(status_dev)
        check(lib().mf_batch_run_device_async(self.h, fp, nF, sp, int(bool(update_coords)), st))

    def synchronize(self):
        check(lib().mf_batch_synchronize(self.h))

    def plan(self):
        p = np.zeros((self.n, 4), np.int32)
        ng = lib().mf_batch_plan(self.h, _iptr(p))
        return ng, p


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ptr(x):
    return None if x is None else (x.data_ptr() if hasattr(x, "data_ptr") else int(x))


def default_train_params(**kw):
    tp = mf_train_params()
    lib().mf_default_train_params(C.byref(tp))
    for k, v in kw.items():
        setattr(tp, k, v)
    return tp


class SurrogateMLP:
    """Ensemble of BF16-MFMA MLPs 10 -> 256 -> 256 -> 256 -> 6 (SiLU) over the
    C ABI (mf_mlp_*).  Inputs: strain [e_xx, e_yy, e_zz, g_yz, g_xz, g_xy]
    of sym(F) - I, node count, fibers per volume, mean fiber length,
    nonlinear fraction; outputs: Cauchy stress (Voigt)."""

    NUM_PARAMS = 135942

    def __init__(self, n_members=4, seed=1):
        h = C.c_void_p()
        check(lib().mf_mlp_create(int(n_members), int(seed), C.byref(h)))
        self.h = h
        self.n_members = int(n_members)

    def close(self):
        if self.h:
            lib().mf_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_normalization(self, in_mean, in_std, out_mean, out_std):
        a = [np.ascontiguousarray(v, dtype=np.float32) for v in (in_mean, in_std, out_mean, out_std)]
        check(lib().mf_mlp_set_normalization(self.h, *[_fptr(v) for v in a]))

    def normalization(self):
        a = [np.zeros(MF_MLP_IN, np.float32), np.zeros(MF_MLP_IN, np.float32),
             np.zeros(MF_MLP_OUT, np.float32), np.zeros(MF_MLP_OUT, np.float32)]
        check(lib().mf_mlp_get_normalization(self.h, *[_fptr(v) for v in a]))
        return a

    def get_weights(self, member=0):
        w = np.zeros(self.NUM_PARAMS, np.float32)
        check(lib().mf_mlp_get_weights(self.h, int(member), _fptr(w)))
        return w

    def set_weights(self, member, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.NUM_PARAMS
        check(lib().mf_mlp_set_weights(self.h, int(member), _fptr(w)))

    def train_step(self, X_dev, Y_dev, n_data, batch, lr, idx_dev=None, return_loss=True):
        loss = C.c_float()
        check(lib().mf_mlp_train_step_device(self.h, _ptr(X_dev), _ptr(Y_dev), int(n_data),
                                             _ptr(idx_dev), int(batch), float(lr),
                                             C.byref(loss) if return_loss else None))
        return loss.value if return_loss else None

    def fit(self, X_dev, Y_dev, n, **kw):
        """Adam epochs on device data; returns the epoch losses (epochs x members)."""
        tp = default_train_params(**kw)
        hist = np.zeros((tp.epochs, self.n_members), np.float32)
        check(lib().mf_mlp_fit_device(self.h, _ptr(X_dev), _ptr(Y_dev), int(n), C.byref(tp),
                                      _fptr(hist)))
        return hist

    def gradient(self, member, X_dev, Y_dev, n):
        g = np.zeros(self.NUM_PARAMS, np.float32)
        loss = C.c_float()
        check(lib().mf_mlp_gradient_device(self.h, int(member), _ptr(X_dev), _ptr(Y_dev),
                                           int(n), _fptr(g), C.byref(loss)))
        return loss.value, g

    def predict_device(self, X_dev, n, Y_dev, J_dev=None, spread_dev=None, member=-1):
        check(lib().mf_mlp_predict_device(self.h, int(member), _ptr(X_dev), int(n),
                                          _ptr(Y_dev), _ptr(J_dev), _ptr(spread_dev)))

    def predict(self, X, jacobian=False, member=-1):
        """Host convenience: returns (Y, spread, J or None)."""
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, MF_MLP_IN)
        n = len(X)
        Xd = DeviceArray.from_host(X)
        Yd = DeviceArray((n, MF_MLP_OUT), np.float32)
        Sd = DeviceArray((n, MF_MLP_OUT), np.float32)
        Jd = DeviceArray((n, 6, 6), np.float32) if jacobian else None
        self.predict_device(Xd, n, Yd, Jd, Sd, member)
        return Yd.numpy(), Sd.numpy(), (Jd.numpy() if jacobian else None)

    def last_predict_ms(self):
        v = C.c_double()
        check(lib().mf_mlp_last_predict_ms(self.h, C.byref(v)))
        return v.value

    def set_estimator(self, in_lo, in_hi, margin=0.0, kappa=1.0, sigma_floor=0.0, tol=1e30):
        e = mf_estimator()
        for i in range(MF_MLP_IN):
            e.in_lo[i] = float(in_lo[i])
            e.in_hi[i] = float(in_hi[i])
        e.margin, e.kappa, e.sigma_floor, e.tol = margin, kappa, sigma_floor, tol
        check(lib().mf_mlp_set_estimator(self.h, C.byref(e)))

    def estimator(self):
        e = mf_estimator()
        check(lib().mf_mlp_get_estimator(self.h, C.byref(e)))
        return e


def surrogate_inputs(analysis, F):
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(-1, 9))
    X = np.zeros((analysis.n, MF_MLP_IN), np.float32)
    check(lib().mf_surrogate_inputs(analysis.h, _dptr(F), F.shape[0], _fptr(X)))
    return X


def surrogate_run(analysis, mlp, F, tangent=False):
    """Surrogate stress (and tangent) with explicit fallback for every RVE of
    a ramp-mode batch at its reference state.  Returns sigma, C (or None),
    status, stats dict."""
    F = np.ascontiguousarray(np.asarray(F, dtype=np.float64).reshape(-1, 9))
    sigma = np.zeros((analysis.n, 6))
    Cm = np.zeros((analysis.n, 6, 6)) if tangent else None
    status = np.zeros(analysis.n, np.int32)
    st = mf_surrogate_stats()
    check(lib().mf_surrogate_run(analysis.h, mlp.h, _dptr(F), F.shape[0], _dptr(sigma),
                                 _dptr(Cm) if tangent else None, _iptr(status), C.byref(st)))
    stats = {k: getattr(st, k) for k, _ in mf_surrogate_stats._fields_}
    return sigma, Cm, status, stats


MF_FO_HEADER_FIELDS, MF_FO_PARAM_FIELDS, MF_FO_SOLVER_FIELDS, MF_FO_SOLVER_INT_FIELDS = 7, 8, 8, 10
MF_FO_FIBER_ONLY, MF_FO_ISOTROPIC_NEOHOOKEAN, MF_FO_SOLVER_EXPLICIT = 1, 3, 1


def fo_scale_conversion(network, fiber_radius, volume_fraction):
    """calcScaleConversion (MultiscaleCoupling.cc:7-22) for one network."""
    v = C.c_double()
    d = network.to_c()
    check(lib().mf_fo_scale_conversion(C.byref(d), float(fiber_radius), float(volume_fraction),
                                       C.byref(v)))
    return v.value


class MultiscaleRVEAdapter:
    """The micro side of MultiscaleRVEAnalysis::run over the batch: per-RVE
    micro_fo_* records in, micro_fo_result (45 doubles) per macro iteration
    and micro_fo_step_result (18 doubles) per macro step out."""

    def __init__(self, networks, hdrs, prms, slvr, slvr_int, params=None):
        n = len(networks)
        self._nets = (mf_network * n)(*[x.to_c() for x in networks])
        self._keep = networks
        h = np.ascontiguousarray(hdrs, dtype=np.int32).reshape(n, MF_FO_HEADER_FIELDS)
        p = np.ascontiguousarray(prms, dtype=np.float64).reshape(n, MF_FO_PARAM_FIELDS)
        s = np.ascontiguousarray(slvr, dtype=np.float64).reshape(n, MF_FO_SOLVER_FIELDS)
        si = np.ascontiguousarray(slvr_int, dtype=np.int32).reshape(n, MF_FO_SOLVER_INT_FIELDS)
        hp = C.c_void_p()
        check(lib().mf_fo_create(self._nets, n, _iptr(h), _dptr(p), _dptr(s), _iptr(si),
                                 C.byref(params) if params is not None else None, C.byref(hp)))
        self.h = hp
        self.n = n

    def close(self):
        if self.h:
            lib().mf_fo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, data, return_status=False):
        d = np.ascontiguousarray(data, dtype=np.float64).reshape(self.n, 9)
        res = np.zeros((self.n, 45))
        st = np.zeros(self.n, np.int32)
        check(lib().mf_fo_run(self.h, _dptr(d), _dptr(res), _iptr(st)))
        return (res, st) if return_status else res

    def step_results(self):
        out = np.zeros((self.n, 18))
        check(lib().mf_fo_step_results(self.h, _dptr(out)))
        return out
