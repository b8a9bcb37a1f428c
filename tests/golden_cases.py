"""Definitions of the golden-vector cases (tests/golden/*.npz).

build(name) returns (Case, FP64 initial state, JSON-able description). The same
description is stored in each fixture, and cases_from_meta() rebuilds the Case
from it, so a test reads everything it needs from the fixture itself.
"""
import numpy as np

from cases import (Case, Plane, TAG_BB, TAG_MEBB, TAG_NODYN, block_boundary_objects, cavity_case, channel_case,
                   doctest_cavity_case, random_state, rest_state, seeded_velocity)

# Uneven block decomposition of the 12 x 10 x 9 multi-block case (inclusive boxes).
MULTIBLOCK_BLOCKS = [
    (0, 4, 0, 3, 0, 8), (0, 4, 4, 9, 0, 3), (0, 4, 4, 9, 4, 8),
    (5, 11, 0, 6, 0, 5), (5, 11, 0, 6, 6, 8), (5, 11, 7, 9, 0, 8),
]

GOLDEN = {
    "periodic_bgk_f64": {"dumps": [1, 20]},
    "periodic_bgk_f32": {"dumps": [1, 20]},
    "periodic_rbgk_f64": {"dumps": [1, 20]},
    "periodic_rbgk_f32": {"dumps": [20]},
    "walls_bb_nodyn_f64": {"dumps": [15]},
    "cavity_zouhe_f64": {"dumps": [50]},
    "cavity_zouhe_f32": {"dumps": [50]},
    "cavity_zouhe_wrapped_f64": {"dumps": [50], "bcstyle": "wrapped"},
    "cavity_regularized_f64": {"dumps": [50]},
    "channel_zouhe_mebb_f64": {"dumps": [30]},
    "channel_zouhe_mebb_f32": {"dumps": [30]},
    "channel_regularized_bb_rbgk_f64": {"dumps": [30]},
    "faces_zouhe_f64": {"dumps": [12]},
    "faces_zouhe_swapped_f64": {"dumps": [12]},
    "faces_regularized_f64": {"dumps": [12]},
    "faces_regularized_swapped_f64": {"dumps": [12]},
    "guo_f64": {"dumps": [11]},
    "guo_f32": {"dumps": [11]},
    "multiblock_mebb_f64": {"dumps": [10], "blocks": MULTIBLOCK_BLOCKS},
    "cavity64_f64": {"dumps": [1000], "summary": True},
    # velocity conditions on every face, edge and corner (setVelocityConditionOnBlockBoundaries)
    "blockbc_regularized_wrapped_f64": {"dumps": [1, 7], "bcstyle": "wrapped"},
    "blockbc_regularized_rbgk_f64": {"dumps": [1, 7]},
    "blockbc_zouhe_f64": {"dumps": [7]},
    "blockbc_zouhe_wrapped_f32": {"dumps": [1, 7], "bcstyle": "wrapped"},
    "blockbc_regularized_f32": {"dumps": [7]},
    "blockbc_outflow_f64": {"dumps": [1, 6]},
    "blockbc_freeslip_f64": {"dumps": [6]},
    "blockbc_normaloutflow_wrapped_f64": {"dumps": [6], "bcstyle": "wrapped"},
    "pressure_neumann_f64": {"dumps": [5]},
    # the jlabos doctest cavities (jlabos/src/testing/test_cylinder.py:85-98), 30 steps
    "doctest_cavity_20x20x20_f64": {"dumps": [30], "summary": True, "bcstyle": "wrapped"},
    "doctest_cavity_39x39x39_f64": {"dumps": [30], "summary": True, "bcstyle": "wrapped"},
    "doctest_cavity_58x58x58_f64": {"dumps": [30], "summary": True, "bcstyle": "wrapped"},
    "doctest_cavity_20x31x42_f64": {"dumps": [30], "summary": True, "bcstyle": "wrapped"},
    "doctest_cavity_20x31x42_f32": {"dumps": [30], "summary": True, "bcstyle": "wrapped"},
    # Palabos's external-moment cycle in place of collideAndStream: BoxRhoBarJfunctional3D,
    # j_x += a (1 + rhoBar), ExternalRhoJcollideAndStream3D (periodic, BounceBack and NoDynamics cells)
    "external_bgk_f64": {"dumps": [1, 6], "external": 2e-4},
    "external_rbgk_f32": {"dumps": [1, 5], "external": 2e-4},
    "external_rbgk_f64": {"dumps": [4], "external": -3e-4},
}


def all_faces(storage, family, swap_kinds=False, seed=5):
    """Velocity planes on the x and z faces, pressure on the y faces (or swapped),
    face interiors only: every (direction, orientation) pair of both kinds."""
    nx, ny, nz = 10, 11, 12
    planes = []
    rng = np.random.RandomState(seed)
    for d, n in enumerate((nx, ny, nz)):
        for ori in (-1, 1):
            lo = [1, 1, 1]
            hi = [nx - 2, ny - 2, nz - 2]
            pos = 0 if ori < 0 else n - 1
            lo[d] = hi[d] = pos
            box = (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
            cells = (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1)
            if (d != 1) != swap_kinds:
                planes.append(Plane(family, "velocity", d, ori, box, 0.03 * (rng.rand(cells, 3) - 0.5)))
            else:
                planes.append(Plane(family, "pressure", d, ori, box, 1.0 + 0.01 * (rng.rand(cells) - 0.5)))
    return Case(nx, ny, nz, storage=storage, omega=1.6, planes=planes)


def walls_case(storage):
    c = Case(8, 9, 10, storage=storage, omega=1.7, periodic=(True, False, False))
    tags = np.zeros(c.shape, np.uint8)
    tags[3:5, 3:6, 4:6] = TAG_BB
    tags[6, 1:3, 1:3] = TAG_NODYN
    c.tags = tags
    return c


def _make(name):
    storage = "f32" if name.endswith("_f32") else "f64"
    base = name.rsplit("_", 1)[0]
    init_seed = None
    if base == "periodic_bgk":
        case, init_seed = Case(8, 9, 10, storage=storage, omega=1.7, periodic=(True, True, True)), 11
    elif base == "periodic_rbgk":
        case, init_seed = Case(8, 9, 10, storage=storage, dynamics="rbgk", omega=1.3,
                               periodic=(True, True, True)), 11
    elif base == "walls_bb_nodyn":
        case, init_seed = walls_case(storage), 13
    elif base in ("cavity_zouhe", "cavity_zouhe_wrapped"):
        case = cavity_case(12, 1 / 0.6, 0.05, storage=storage)
    elif base == "cavity_regularized":
        case = cavity_case(12, 1 / 0.6, 0.05, storage=storage, family="regularized")
    elif base == "channel_zouhe_mebb":
        case = channel_case(16, 10, 11, 1.8, 0.05, storage=storage, sphere=((6, 5, 5), 2.2))
    elif base == "channel_regularized_bb_rbgk":
        case = channel_case(16, 10, 11, 1.8, 0.05, storage=storage, family="regularized",
                            sphere=((6, 5, 5), 2.2), sphere_tag=TAG_BB, dynamics="rbgk")
    elif base.startswith("faces_"):
        family = "zouhe" if "zouhe" in base else "regularized"
        case, init_seed = all_faces(storage, family, swap_kinds=base.endswith("swapped")), 7
    elif base == "guo":
        case, init_seed = Case(8, 9, 10, storage=storage, omega=1.7, periodic=(True, True, True),
                               force=(1e-4, -5e-5, 2e-5)), 11
    elif base == "multiblock_mebb":
        c = Case(12, 10, 9, storage=storage, omega=1.6, periodic=(True, False, False))
        from cases import sphere_tags
        c.tags = sphere_tags(c.shape, (5.5, 4.7, 4.2), 2.1, TAG_MEBB)
        case, init_seed = c, 17
    elif base == "cavity64":
        case = cavity_case(64, 1 / 0.6, 0.05, storage=storage)
    elif base.startswith("blockbc_"):
        family = "zouhe" if "zouhe" in base else "regularized"
        bc_type = {"outflow": "outflow", "freeslip": "freeslip", "normaloutflow": "normalOutflow"}.get(
            base.split("_")[1], "dirichlet")
        shape = (8, 9, 10)
        case = Case(*shape, storage=storage, dynamics="rbgk" if "rbgk" in base else "bgk", omega=1.6,
                    planes=block_boundary_objects(shape, family, bc_type=bc_type, velocity=seeded_velocity(5)))
        init_seed = 3
    elif base == "pressure_neumann":
        shape = (8, 9, 10)
        case = Case(*shape, storage=storage, omega=1.7,
                    planes=block_boundary_objects(shape, "zouhe", bc_type="neumann", kind="pressure"))
        init_seed = 3
    elif base.startswith("external_"):
        case = Case(8, 9, 10, storage=storage, dynamics="rbgk" if "rbgk" in base else "bgk", omega=1.6,
                    periodic=(True, True, True))
        tags = np.zeros(case.shape, np.uint8)
        tags[3:5, 3:6, 4:6] = TAG_BB
        tags[6, 1:3, 1:3] = TAG_NODYN
        case.tags = tags
        init_seed = 19
    elif base.startswith("doctest_cavity_"):
        nx, ny, nz = (int(v) for v in base.split("_")[2].split("x"))
        case = doctest_cavity_case(nx, ny, nz, storage=storage)
    else:
        raise KeyError(name)
    return case, init_seed


def build(name):
    case, seed = _make(name)
    init = random_state(case, seed) if seed is not None else rest_state(case)
    meta = {"name": name, "seed": seed}
    return case, init, meta


def load(name, directory):
    """Rebuilds (case, init, fixture dict) for a stored fixture."""
    import hashlib
    import json
    import os
    data = dict(np.load(os.path.join(directory, name + ".npz"), allow_pickle=False))
    meta = json.loads(str(data["case_json"]))
    case, init, _ = build(meta["name"])
    digest = hashlib.sha256(np.ascontiguousarray(init).tobytes()).hexdigest()
    if digest != str(data["init_sha256"]):
        raise AssertionError(f"{name}: the regenerated initial state differs from the one Palabos was run on")
    return case, init, data
