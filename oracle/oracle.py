"""ctypes wrapper of the oracle restatement (oracle/liboracle.so).

Test infrastructure only: tests/, the smoke check and the benchmark's
CPU-baseline leg use it as the checker. The product never imports it.

A lattice case is any object with the attributes of ``tests/cases.py:Case``:
nx, ny, nz, storage ('f64' | 'f32' | 'f16' | 'bf16'), dynamics ('bgk' | 'rbgk'),
omega, periodic (3 bools), tags (uint8 [nx,ny,nz] or None), planes (list of
Plane), force (3 floats), alpha (float64 [nx,ny,nz] or None).

Populations are numpy arrays in Palabos's static-data order, [nx][ny][nz][19]
of shifted populations f-bar (``BlockLatticeDataTransfer3D::send``,
atomicBlock/blockLattice3D.hh:650-690), in the storage dtype
(float64, float32, or uint16 bit patterns for f16/bf16).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

STORAGE_CODES = {"f64": 0, "f32": 1, "f16": 2, "bf16": 3}
STORAGE_DTYPES = {"f64": np.float64, "f32": np.float32, "f16": np.uint16, "bf16": np.uint16}
DYNAMICS_CODES = {"bgk": 0, "rbgk": 1}
FAMILY_CODES = {"zouhe": 0, "regularized": 1}
KIND_CODES = {"velocity": 0, "pressure": 1}
SHAPE_CODES = {"plane": 0, "edge": 1, "corner": 2}
# boundary::BcType (boundaryCondition/boundaryCondition.h:40)
BC_TYPE_CODES = {"dirichlet": 0, "neumann": 1, "freeslip": 2, "density": 3, "outflow": 4, "normalOutflow": 5}


class _Plane(ctypes.Structure):
    _fields_ = [("family", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("dir", ctypes.c_int32), ("ori", ctypes.c_int32),
                ("box", ctypes.c_int32 * 6), ("values", ctypes.c_void_p),
                ("shape", ctypes.c_int32), ("normal", ctypes.c_int32 * 3), ("bc_type", ctypes.c_int32)]


class _Desc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("storage", ctypes.c_int32), ("dynamics", ctypes.c_int32),
                ("periodic", ctypes.c_int32 * 3), ("omega", ctypes.c_double),
                ("tags", ctypes.c_void_p), ("nplanes", ctypes.c_int32),
                ("planes", ctypes.c_void_p), ("force", ctypes.c_double * 3),
                ("alpha", ctypes.c_void_p), ("bc_wrapped", ctypes.c_int32)]


def build():
    """Compiles oracle/liboracle.so with the committed Makefile."""
    root = os.path.dirname(_HERE)
    subprocess.run(["make", "-s", "-f", "oracle/Makefile", "all"], cwd=root, check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_run.argtypes = [ctypes.POINTER(_Desc), ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_void_p]
        _lib.oracle_run.restype = ctypes.c_int
        _lib.oracle_run_ex.argtypes = [ctypes.POINTER(_Desc), ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_int]
        _lib.oracle_run_ex.restype = ctypes.c_int
        _lib.oracle_complete.argtypes = [ctypes.POINTER(_Desc), ctypes.c_void_p]
        _lib.oracle_complete.restype = ctypes.c_int
        _lib.oracle_equilibrium.argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p]
        _lib.oracle_objective_dual.argtypes = [ctypes.POINTER(_Desc), ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        _lib.oracle_objective_dual.restype = ctypes.c_int
        _lib.oracle_fields.argtypes = [ctypes.POINTER(_Desc), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.oracle_fields.restype = ctypes.c_int
        _lib.oracle_rhobar_j.argtypes = [ctypes.POINTER(_Desc), ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p,
                                         ctypes.c_void_p]
        _lib.oracle_rhobar_j.restype = ctypes.c_int
        _lib.oracle_external_step.argtypes = [ctypes.POINTER(_Desc), ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p]
        _lib.oracle_external_step.restype = ctypes.c_int
        _lib.oracle_exact_rates.argtypes = [ctypes.c_float, ctypes.c_void_p]
        _lib.oracle_exact_rates.restype = None
        for name in ("oracle_f32_to_f16", "oracle_f16_to_f32", "oracle_f32_to_bf16",
                     "oracle_bf16_to_f32"):
            getattr(_lib, name).argtypes = [ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p]
    return _lib


class _Keep:
    """Holds the numpy buffers a descriptor points into."""

    def __init__(self):
        self.refs = []

    def ptr(self, arr):
        self.refs.append(arr)
        return arr.ctypes.data


def _desc(case, keep):
    d = _Desc()
    d.nx, d.ny, d.nz = case.nx, case.ny, case.nz
    d.storage = STORAGE_CODES[case.storage]
    d.dynamics = DYNAMICS_CODES[case.dynamics]
    for a in range(3):
        d.periodic[a] = int(bool(case.periodic[a]))
    d.omega = float(case.omega)
    if case.tags is not None:
        tags = np.ascontiguousarray(case.tags, dtype=np.uint8)
        assert tags.shape == (case.nx, case.ny, case.nz)
        d.tags = keep.ptr(tags)
    planes = list(case.planes or [])
    d.nplanes = len(planes)
    if planes:
        arr = (_Plane * len(planes))()
        for i, p in enumerate(planes):
            arr[i].family = FAMILY_CODES[p.family]
            arr[i].kind = KIND_CODES[p.kind]
            arr[i].dir = p.dir
            arr[i].ori = p.ori
            for k in range(6):
                arr[i].box[k] = p.box[k]
            vals = np.ascontiguousarray(p.values, dtype=np.float64)
            arr[i].values = keep.ptr(vals)
            arr[i].shape = SHAPE_CODES[getattr(p, "shape", "plane")]
            for a in range(3):
                arr[i].normal[a] = int(getattr(p, "normal", (0, 0, 0))[a])
            arr[i].bc_type = BC_TYPE_CODES[getattr(p, "bc_type", "dirichlet")]
        keep.refs.append(arr)
        d.planes = ctypes.addressof(arr)
    for a in range(3):
        d.force[a] = float(case.force[a]) if case.force is not None else 0.0
    if case.alpha is not None:
        alpha = np.ascontiguousarray(case.alpha, dtype=np.float64)
        assert alpha.shape == (case.nx, case.ny, case.nz)
        d.alpha = keep.ptr(alpha)
    return d


def run(case, f, nsteps, initialize=False, wrapped=False, snapshot=False):
    """Returns (populations after nsteps, stats dict). `f` is not modified.
    initialize: one pass of the data processors first (Palabos's
    MultiBlock3D::initialize). wrapped: the boundary objects come from a
    wrapped factory (processors read plane cells completed). snapshot: the plane
    cells are returned completed with the boundary values of the last step, as a
    wrapped factory leaves them (the values Neumann-type boundaries copied are
    lost between calls, so a snapshot is taken in the same call)."""
    dtype = STORAGE_DTYPES[case.storage]
    state = np.array(f, dtype=dtype, order="C", copy=True)
    assert state.shape == (case.nx, case.ny, case.nz, 19)
    keep = _Keep()
    d = _desc(case, keep)
    d.bc_wrapped = int(bool(wrapped))
    stats = np.zeros(10, dtype=np.float64)
    flags = (1 if initialize else 0) | (2 if snapshot else 0)
    rc = lib().oracle_run_ex(ctypes.byref(d), state.ctypes.data, int(nsteps), stats.ctypes.data, flags)
    if rc != 0:
        raise ValueError(f"oracle_run rejected the case ({rc})")
    return state, {"avg_rhobar": stats[0], "avg_usqr": stats[1], "max_usqr": stats[2],
                   "force": stats[3:6].copy(), "count": int(stats[6]),
                   "porous": stats[7:10].copy()}


def fields(case, f, solid_rho=(1.0, 1.0, 1.0)):
    """Analysis fields [nx, ny, nz, 7] of the state `f` (computeDensity,
    computeRhoBar, computeVelocity x y z, computeKineticEnergy,
    computeVelocityNorm) in the storage's arithmetic type, widened to float64;
    solid_rho: density parameters of BounceBack, NoDynamics,
    MomentumExchangeBounceBack. Lattices with planes are not covered."""
    dtype = STORAGE_DTYPES[case.storage]
    state = np.ascontiguousarray(f, dtype=dtype)
    assert state.shape == (case.nx, case.ny, case.nz, 19)
    keep = _Keep()
    d = _desc(case, keep)
    sr = np.asarray(solid_rho, dtype=np.float64)
    out = np.empty((case.nx, case.ny, case.nz, 7), dtype=np.float64)
    rc = lib().oracle_fields(ctypes.byref(d), state.ctypes.data, sr.ctypes.data, out.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle_fields rejected the case ({rc})")
    return out


def complete(case, f):
    """Applies the plane-boundary completion (wrapped-processor snapshot convention)."""
    dtype = STORAGE_DTYPES[case.storage]
    state = np.array(f, dtype=dtype, order="C", copy=True)
    keep = _Keep()
    d = _desc(case, keep)
    if lib().oracle_complete(ctypes.byref(d), state.ctypes.data) != 0:
        raise ValueError("oracle_complete rejected the case")
    return state


def objective_dual(case, f0, nsteps, xin, xout, w_dp, w_drag, seeds, w_flux=0.0, first_step=0):
    """J = 1/(T - t0) sum_{t > t0} j(f^t) of the adjoint objective and its exact
    derivatives (forward-mode dual numbers through the oracle) with respect to the
    porosity of 16 seeded cells (flat cell indices) and to the velocity-plane scale.
    Returns (J, dJ/dalpha[16], dJ/ds)."""
    f0 = np.ascontiguousarray(f0, dtype=np.float64)
    seeds = np.ascontiguousarray(seeds, dtype=np.int64)
    assert seeds.shape == (16,)
    keep = _Keep()
    d = _desc(case, keep)
    out = np.zeros(18)
    rc = lib().oracle_objective_dual(ctypes.byref(d), f0.ctypes.data, int(nsteps), int(xin), int(xout),
                                     float(w_dp), float(w_drag), float(w_flux), int(first_step),
                                     seeds.ctypes.data, out.ctypes.data)
    if rc != 0:
        raise ValueError("oracle_objective_dual rejected the case")
    return out[0], out[1:17].copy(), out[17]


def equilibrium(rho, u):
    """FP64 bulk equilibrium f-bar for per-cell rho [...] and u [...,3]."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.empty(rho.shape + (19,), dtype=np.float64)
    lib().oracle_equilibrium(rho.size, rho.ctypes.data, u.ctypes.data, out.ctypes.data)
    return out


def rhobar_j(case, f, bb_rho=1.0):
    """BoxRhoBarJfunctional3D of the whole lattice on the storage-typed state `f`:
    (rhoBar [nx, ny, nz], j [nx, ny, nz, 3]) in the arithmetic type, widened to
    float64; bb_rho is the BounceBack cells' density parameter."""
    dtype = STORAGE_DTYPES[case.storage]
    state = np.ascontiguousarray(f, dtype=dtype)
    keep = _Keep()
    d = _desc(case, keep)
    rho_bar = np.empty(case.shape, dtype=np.float64)
    j = np.empty(case.shape + (3,), dtype=np.float64)
    rc = lib().oracle_rhobar_j(ctypes.byref(d), state.ctypes.data, float(bb_rho), rho_bar.ctypes.data, j.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle_rhobar_j rejected the case ({rc})")
    return rho_bar, j


def external_step(case, f, rho_bar, j):
    """One ExternalRhoJcollideAndStream3D step of the storage-typed state `f`
    with the field (rho_bar, j; rounded to the arithmetic type); returns the new
    state."""
    dtype = STORAGE_DTYPES[case.storage]
    state = np.array(f, dtype=dtype, copy=True, order="C")
    rb = np.ascontiguousarray(rho_bar, dtype=np.float64)
    jj = np.ascontiguousarray(j, dtype=np.float64)
    assert rb.shape == case.shape and jj.shape == case.shape + (3,)
    keep = _Keep()
    d = _desc(case, keep)
    rc = lib().oracle_external_step(ctypes.byref(d), state.ctypes.data, rb.ctypes.data, jj.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle_external_step rejected the case ({rc})")
    return state


def exact_rates(omega):
    """The 16-bit storage modes' BGK constants for `omega` (FP32):
    [1 - omega', t0 omega', t1 omega', t4 omega', omega']."""
    out = np.empty(5, dtype=np.float32)
    lib().oracle_exact_rates(ctypes.c_float(omega), out.ctypes.data)
    return out


def _conv(name, arr, out_dtype):
    arr = np.ascontiguousarray(arr)
    out = np.empty(arr.shape, dtype=out_dtype)
    getattr(lib(), name)(arr.size, arr.ctypes.data, out.ctypes.data)
    return out


def f32_to_f16(a):
    return _conv("oracle_f32_to_f16", np.asarray(a, np.float32), np.uint16)


def f16_to_f32(a):
    return _conv("oracle_f16_to_f32", np.asarray(a, np.uint16), np.float32)


def f32_to_bf16(a):
    return _conv("oracle_f32_to_bf16", np.asarray(a, np.float32), np.uint16)


def bf16_to_f32(a):
    return _conv("oracle_bf16_to_f32", np.asarray(a, np.uint16), np.float32)


def to_storage(f64, storage):
    """Converts FP64 f-bar to the storage representation (RNE)."""
    if storage == "f64":
        return np.ascontiguousarray(f64, dtype=np.float64)
    if storage == "f32":
        return np.ascontiguousarray(f64, dtype=np.float32)
    f32 = np.ascontiguousarray(f64, dtype=np.float32)
    return f32_to_f16(f32) if storage == "f16" else f32_to_bf16(f32)


def from_storage(s, storage):
    """Converts storage-typed f-bar to float64."""
    if storage in ("f64", "f32"):
        return np.asarray(s, dtype=np.float64)
    f32 = f16_to_f32(s) if storage == "f16" else bf16_to_f32(s)
    return f32.astype(np.float64)
