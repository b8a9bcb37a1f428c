"""ctypes binding of the C ABI declared in include/palabos_amd.h.

The library is loaded from the package directory (palabos_amd/libpalabos_amd.so,
built by palabos_amd/_build.py). There is no fallback: if the library is
missing or fails to load, NativeLibraryError is raised.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpalabos_amd.so")

OK, EINVAL, ENOMEM, EHIP, ESTATE = 0, -1, -2, -3, -4
F64, F32, F16, BF16 = 0, 1, 2, 3
BGK, RBGK = 0, 1
TAG_BULK, TAG_BOUNCE_BACK, TAG_NO_DYNAMICS, TAG_MOMENTUM_EXCHANGE, TAG_WALL_LINK = 0, 1, 2, 3, 4
ZOUHE, REGULARIZED = 0, 1
WRAPPED = 0x100
KEEP_BLOCKS = 1           # plbamd_multiblock_create_ex: keep the given blocks
BC_TYPES = {"dirichlet": 0, "neumann": 1, "freeslip": 2, "density": 3, "outflow": 4, "normalOutflow": 5}
VELOCITY, PRESSURE = 0, 1
NUM_STATS = 10

STORAGE_CODES = {"f64": F64, "f32": F32, "f16": F16, "bf16": BF16}
STORAGE_NUMPY = {"f64": np.float64, "f32": np.float32, "f16": np.uint16, "bf16": np.uint16}

# Every symbol include/palabos_amd.h declares.
EXPORTS = (
    "plbamd_init", "plbamd_last_error", "plbamd_lattice_create", "plbamd_multiblock_create",
    "plbamd_lattice_destroy",
    "plbamd_set_tags", "plbamd_add_plane", "plbamd_set_plane_values", "plbamd_set_porosity",
    "plbamd_upload", "plbamd_download", "plbamd_initialize_at_equilibrium",
    "plbamd_collide_and_stream", "plbamd_collide_and_stream_timed", "plbamd_toggle_statistics",
    "plbamd_get_statistics", "plbamd_compute_macroscopic", "plbamd_reduce_moments",
    "plbamd_adjoint_gradient", "plbamd_set_adjoint_record_bytes", "plbamd_get_adjoint_record_info", "plbamd_get_time", "plbamd_save_binary_block", "plbamd_load_binary_block",
    "plbamd_compute_fields", "plbamd_set_solid_density", "plbamd_reduce_moments_noalloc",
    "plbamd_add_velocity_edge", "plbamd_add_velocity_corner", "plbamd_set_boundary_type",
    "plbamd_execute_processors", "plbamd_bandwidth_probe", "plbamd_multiblock_create_ex",
    "plbamd_get_block_count", "plbamd_compute_rhobar_j", "plbamd_get_rhobar_j", "plbamd_set_rhobar_j",
    "plbamd_external_collide_and_stream",
)


class NativeLibraryError(RuntimeError):
    pass


class PlbError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"palabos_amd error {code}: {message}")
        self.code = code


class Box(ctypes.Structure):
    _fields_ = [("x0", ctypes.c_int32), ("x1", ctypes.c_int32), ("y0", ctypes.c_int32),
                ("y1", ctypes.c_int32), ("z0", ctypes.c_int32), ("z1", ctypes.c_int32)]

    @classmethod
    def of(cls, b):
        return cls(*[int(v) for v in b])

    def cells(self):
        return (self.x1 - self.x0 + 1) * (self.y1 - self.y0 + 1) * (self.z1 - self.z0 + 1)

    def shape(self):
        return (self.x1 - self.x0 + 1, self.y1 - self.y0 + 1, self.z1 - self.z0 + 1)


class LatticeDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("storage", ctypes.c_int32), ("dynamics", ctypes.c_int32),
                ("periodic", ctypes.c_int32 * 3), ("omega", ctypes.c_double),
                ("force", ctypes.c_double * 3)]


class Objective(ctypes.Structure):
    _fields_ = [("inlet_x", ctypes.c_int32), ("outlet_x", ctypes.c_int32),
                ("weight_pressure_drop", ctypes.c_double), ("weight_drag", ctypes.c_double),
                ("weight_flux", ctypes.c_double), ("first_step", ctypes.c_int32)]


_lib = None


def lib():
    """Loads the native library once; raises NativeLibraryError if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    try:
        L = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise NativeLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    vp, i32, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
    pbox = ctypes.POINTER(Box)
    sig = {
        "plbamd_init": [i32],
        "plbamd_lattice_create": [ctypes.POINTER(LatticeDesc), ctypes.POINTER(ctypes.c_int)],
        "plbamd_multiblock_create": [ctypes.POINTER(LatticeDesc), pbox, i32, ctypes.POINTER(ctypes.c_int)],
        "plbamd_multiblock_create_ex": [ctypes.POINTER(LatticeDesc), pbox, i32, i32, ctypes.POINTER(ctypes.c_int)],
        "plbamd_lattice_destroy": [i32],
        "plbamd_get_block_count": [i32, ctypes.POINTER(ctypes.c_int)],
        "plbamd_compute_rhobar_j": [i32, pbox],
        "plbamd_get_rhobar_j": [i32, pbox, vp, vp],
        "plbamd_set_rhobar_j": [i32, pbox, vp, vp],
        "plbamd_external_collide_and_stream": [i32, i32],
        "plbamd_set_tags": [i32, vp],
        "plbamd_add_plane": [i32, i32, i32, i32, i32, pbox, ctypes.POINTER(ctypes.c_int)],
        "plbamd_set_plane_values": [i32, i32, vp, sz],
        "plbamd_set_porosity": [i32, vp],
        "plbamd_upload": [i32, pbox, i32, vp, sz],
        "plbamd_download": [i32, pbox, i32, vp, sz],
        "plbamd_initialize_at_equilibrium": [i32, pbox, vp, vp],
        "plbamd_collide_and_stream": [i32, i32],
        "plbamd_collide_and_stream_timed": [i32, i32, ctypes.POINTER(ctypes.c_double)],
        "plbamd_toggle_statistics": [i32, i32],
        "plbamd_get_statistics": [i32, vp],
        "plbamd_compute_macroscopic": [i32, pbox, vp, vp],
        "plbamd_reduce_moments": [i32, pbox, vp],
        "plbamd_reduce_moments_noalloc": [i32, pbox, vp],
        "plbamd_add_velocity_edge": [i32, i32, i32, i32, pbox, ctypes.POINTER(ctypes.c_int)],
        "plbamd_add_velocity_corner": [i32, i32, i32, i32, i32, i32, i32, ctypes.POINTER(ctypes.c_int)],
        "plbamd_set_boundary_type": [i32, i32, i32],
        "plbamd_execute_processors": [i32],
        "plbamd_compute_fields": [i32, pbox, vp],
        "plbamd_set_solid_density": [i32, i32, ctypes.c_double],
        "plbamd_get_time": [i32, ctypes.POINTER(ctypes.c_int64)],
        "plbamd_adjoint_gradient": [i32, i32, i32, ctypes.POINTER(Objective), vp, vp, vp, vp],
        "plbamd_set_adjoint_record_bytes": [i32, ctypes.c_int64],
        "plbamd_get_adjoint_record_info": [i32, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)],
        "plbamd_save_binary_block": [i32, ctypes.c_char_p, i32, i32, i32],
        "plbamd_load_binary_block": [i32, ctypes.c_char_p, i32, i32, i32],
        "plbamd_bandwidth_probe": [i32, ctypes.c_uint64, i32, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.plbamd_last_error.argtypes = []
    L.plbamd_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc):
    if rc != OK:
        msg = lib().plbamd_last_error().decode(errors="replace")
        raise PlbError(rc, msg)
    return rc


_initialised = False


PROBE_PATTERNS = {"copy": 0, "slots": 1, "slots4": 2, "slots16": 3}


def bandwidth_probe(pattern, nbytes, repeats=3):
    """Streaming rate of the device for `pattern` ('copy': 16-byte copy into a
    second buffer; 'slots', 'slots4', 'slots16': 19 slots read and written back
    in place in 8-, 4- or 16-byte elements per thread) over `nbytes` moved per
    launch: (shortest launch ms, bytes moved, GB/s)."""
    init()
    ms, moved = ctypes.c_double(), ctypes.c_double()
    check(lib().plbamd_bandwidth_probe(PROBE_PATTERNS[pattern], int(nbytes), int(repeats), ctypes.byref(ms),
                                       ctypes.byref(moved)))
    return ms.value, moved.value, moved.value / (ms.value * 1e-3) / 1e9


def init(device=0):
    global _initialised
    if not _initialised:
        check(lib().plbamd_init(int(device)))
        _initialised = True
