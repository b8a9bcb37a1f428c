"""Lattice cases shared by the parity tests and the golden-vector generator.

A Case describes one lattice (size, precision, dynamics, periodicity, per-cell
tags, boundary planes, body force, porosity). The same Case drives the oracle
(oracle/oracle.py), the HIP engine (palabos_amd.Block3D) and, for golden
vectors, the compiled Palabos reference (oracle/ref_driver.cpp).
"""
import dataclasses
from typing import List, Optional, Tuple

import numpy as np

from palabos_amd.setups import poiseuille_profile  # noqa: F401  (re-exported for the tests)
from palabos_amd.setups import tile_blocks as uneven_blocks  # noqa: F401
from palabos_amd.setups import uneven_blocks as seeded_uneven_blocks  # noqa: F401

TAG_BULK, TAG_BB, TAG_NODYN, TAG_MEBB, TAG_WALL = 0, 1, 2, 3, 4


@dataclasses.dataclass
class Plane:
    family: str            # 'zouhe' | 'regularized'
    kind: str              # 'velocity' | 'pressure'
    dir: int
    ori: int
    box: Tuple[int, int, int, int, int, int]
    values: np.ndarray     # velocity: [cells, 3]; pressure: [cells] (rho)
    shape: str = "plane"   # 'plane' | 'edge' (dir = the edge's axis) | 'corner'
    normal: Tuple[int, int, int] = (0, 0, 0)   # edges / corners: outward normal
    bc_type: str = "dirichlet"                 # boundary::BcType name

    def cells(self):
        x0, x1, y0, y1, z0, z1 = self.box
        return (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1)


@dataclasses.dataclass
class Case:
    nx: int
    ny: int
    nz: int
    storage: str = "f64"
    dynamics: str = "bgk"
    omega: float = 1.0
    periodic: Tuple[bool, bool, bool] = (False, False, False)
    tags: Optional[np.ndarray] = None
    planes: List[Plane] = dataclasses.field(default_factory=list)
    force: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    alpha: Optional[np.ndarray] = None

    def with_storage(self, storage):
        return dataclasses.replace(self, storage=storage)

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)


# ---------------------------------------------------------------------------
# Deterministic pseudo-random fields (splitmix64 of index and seed), identical
# on every platform and numpy version.

def _splitmix(idx, seed):
    with np.errstate(over="ignore"):
        z = (idx.astype(np.uint64) + np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)
             + np.uint64(0x632BE59BD9B4E019))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(shape, seed):
    """Uniform values in [-1, 1), from 53-bit integers."""
    n = int(np.prod(shape))
    z = _splitmix(np.arange(n, dtype=np.uint64), seed)
    return ((z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53) * 2.0 - 1.0).reshape(shape)


def equilibrium(rho, u):
    """FP64 shifted equilibrium f-bar (same formula as the oracle and Palabos)."""
    from oracle import oracle
    return oracle.equilibrium(rho, u)


def random_state(case, seed, u_amp=0.05, rho_amp=0.02, noise=5e-4):
    """A non-equilibrium FP64 state: equilibrium of random rho/u plus noise."""
    shape = case.shape
    rho = 1.0 + rho_amp * uniform(shape, seed)
    u = u_amp * uniform(shape + (3,), seed + 1)
    f = equilibrium(rho, u)
    f = f + noise * uniform(shape + (19,), seed + 2)
    return np.ascontiguousarray(f)


def rest_state(case):
    return np.zeros(case.shape + (19,), dtype=np.float64)


# ---------------------------------------------------------------------------
# Geometry helpers

def sphere_tags(shape, center, radius, tag):
    x, y, z = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    r2 = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
    tags = np.zeros(shape, dtype=np.uint8)
    tags[r2 <= radius * radius] = tag
    return tags


def channel_case(nx, ny, nz, omega, u_max, storage="f64", family="zouhe", sphere=None,
                 sphere_tag=TAG_MEBB, dynamics="bgk"):
    """Velocity inlet (x=0) and pressure outlet (x=nx-1) on the face interiors;
    the side walls are the domain edges (same-step reflection)."""
    planes = []
    prof = poiseuille_profile(ny, nz, u_max)
    vel = np.zeros(((ny - 2) * (nz - 2), 3))
    vel[:, 0] = prof.reshape(-1)
    planes.append(Plane(family, "velocity", 0, -1, (0, 0, 1, ny - 2, 1, nz - 2), vel))
    planes.append(Plane(family, "pressure", 0, 1, (nx - 1, nx - 1, 1, ny - 2, 1, nz - 2),
                        np.ones((ny - 2) * (nz - 2))))
    tags = None
    if sphere is not None:
        center, radius = sphere
        tags = sphere_tags((nx, ny, nz), center, radius, sphere_tag)
    return Case(nx, ny, nz, storage=storage, dynamics=dynamics, omega=omega, tags=tags,
                planes=planes)


def cavity_case(n, omega, u_lid, storage="f64", family="zouhe", ny=None, nz=None, dynamics="bgk"):
    """Lid-driven cavity: Zou-He (or regularized) velocity lid on the top z face
    interior Box3D(1, nx-2, 1, ny-2, nz-1, nz-1); every other wall is the domain edge."""
    nx = n
    ny = ny or n
    nz = nz or n
    box = (1, nx - 2, 1, ny - 2, nz - 1, nz - 1)
    cells = (nx - 2) * (ny - 2)
    vel = np.zeros((cells, 3))
    vel[:, 0] = u_lid
    return Case(nx, ny, nz, storage=storage, dynamics=dynamics, omega=omega,
                planes=[Plane(family, "velocity", 2, 1, box, vel)])


def seeded_velocity(seed, amp=0.02):
    """A deterministic velocity per cell (x, y, z) -> (ux, uy, uz) in [-amp, amp)."""
    def f(x, y, z):
        return tuple(amp * uniform((3,), seed * 100003 + (x * 1000 + y) * 1000 + z))
    return f


def doctest_cavity_case(nx, ny, nz, storage="f64"):
    """The 3D regression case of the reference's jlabos doctest
    (jlabos/src/testing/test_cylinder.py:78-130): BGK with omega = 1/(3 nu + 1/2),
    nu = 0.02 ny / 100; regularized velocity conditions of the wrapped factory
    (createLocalBoundaryCondition3D) on every face, edge and corner, zero velocity
    except the lid Box3D(1, nx-2, 1, ny-2, nz-1, nz-1) with u = (0.02, 0, 0); the
    lattice starts at rest. The doctest runs 30 steps."""
    u_max, re = 0.02, 100.0
    nu = u_max * ny / re
    tau = 3 * nu + 0.5

    def vel(x, y, z):
        return (u_max, 0.0, 0.0) if (z == nz - 1 and 1 <= x <= nx - 2 and 1 <= y <= ny - 2) else (0.0, 0.0, 0.0)

    planes = block_boundary_objects((nx, ny, nz), "regularized", velocity=vel)
    return Case(nx, ny, nz, storage=storage, omega=1.0 / tau, planes=planes)


def block_boundary_objects(shape, family, bc_type="dirichlet", velocity=None, kind="velocity"):
    """The boundary objects OnLatticeBoundaryCondition3D::set{Velocity,Pressure}
    ConditionOnBlockBoundaries installs on a whole nx x ny x nz lattice
    (boundaryCondition3D.hh:300-400, BlockSurface3D boxes of core/blockSurface3D.cpp:44-198),
    in its order: the six face interiors (0N 0P 1N 1P 2N 2P), then, for velocity
    conditions, the twelve edges (plane 0 NN NP PN PP, plane 1 ..., plane 2 ...;
    normal1 along (plane+1)%3, normal2 along (plane+2)%3) and the eight corners
    (NNN NNP NPN NPP PNN PNP PPN PPP). velocity(x, y, z) -> (ux, uy, uz) gives the
    stored velocity of each cell (default 0); pressure faces store rho = 1."""
    n = shape
    lo_hi = [(0, n[a] - 1) for a in range(3)]
    vel = velocity or (lambda x, y, z: (0.0, 0.0, 0.0))

    def cells(box):
        x0, x1, y0, y1, z0, z1 = box
        return [(x, y, z) for x in range(x0, x1 + 1) for y in range(y0, y1 + 1) for z in range(z0, z1 + 1)]

    def values(box):
        return np.array([vel(*c) for c in cells(box)], dtype=np.float64).reshape(-1, 3)

    out = []
    for d in range(3):
        for ori in (-1, 1):
            lo = [1, 1, 1]
            hi = [n[0] - 2, n[1] - 2, n[2] - 2]
            lo[d] = hi[d] = lo_hi[d][0] if ori < 0 else lo_hi[d][1]
            box = (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
            if kind == "velocity":
                out.append(Plane(family, "velocity", d, ori, box, values(box), bc_type=bc_type))
            else:
                out.append(Plane(family, "pressure", d, ori, box, np.ones(len(cells(box))), bc_type=bc_type))
    if kind != "velocity":
        return out
    for plane in range(3):
        d1, d2 = (plane + 1) % 3, (plane + 2) % 3
        for n1 in (-1, 1):
            for n2 in (-1, 1):
                lo = [0, 0, 0]
                hi = [0, 0, 0]
                lo[plane], hi[plane] = 1, n[plane] - 2
                lo[d1] = hi[d1] = 0 if n1 < 0 else n[d1] - 1
                lo[d2] = hi[d2] = 0 if n2 < 0 else n[d2] - 1
                normal = [0, 0, 0]
                normal[d1], normal[d2] = n1, n2
                box = (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])
                out.append(Plane(family, "velocity", plane, 0, box, values(box), shape="edge",
                                 normal=tuple(normal), bc_type=bc_type))
    for xn in (-1, 1):
        for yn in (-1, 1):
            for zn in (-1, 1):
                x = 0 if xn < 0 else n[0] - 1
                y = 0 if yn < 0 else n[1] - 1
                z = 0 if zn < 0 else n[2] - 1
                box = (x, x, y, y, z, z)
                out.append(Plane(family, "velocity", 0, 0, box, values(box), shape="corner",
                                 normal=(xn, yn, zn), bc_type=bc_type))
    return out


# ---------------------------------------------------------------------------
# Running a case on the HIP engine

def gpu_lattice(case, wrapped=False):
    import palabos_amd as plb
    dyn = plb.BGK(case.omega) if case.dynamics == "bgk" else plb.RegularizedBGK(case.omega)
    force = case.force if any(case.force) else None
    lat = plb.Block3D(case.nx, case.ny, case.nz, plb.D3Q19, dyn, storage=case.storage,
                      periodic=case.periodic, force=force)
    if case.tags is not None:
        lat.setTags(case.tags)
    for p in case.planes:
        bt = getattr(p, "bc_type", "dirichlet")
        shape = getattr(p, "shape", "plane")
        if shape == "edge":
            d1, d2 = (p.dir + 1) % 3, (p.dir + 2) % 3
            pid = lat.addExternalVelocityEdge(p.dir, p.normal[d1], p.normal[d2], p.box, bc_type=bt)
            lat.setBoundaryVelocity(pid, p.values)
        elif shape == "corner":
            pid = lat.addExternalVelocityCorner(*p.normal, p.box[0], p.box[2], p.box[4], bc_type=bt)
            lat.setBoundaryVelocity(pid, p.values)
        elif p.kind == "velocity":
            pid = lat.addVelocityBoundary(p.dir, p.ori, p.box, family=p.family, wrapped=wrapped, bc_type=bt)
            lat.setBoundaryVelocity(pid, p.values)
        else:
            pid = lat.addPressureBoundary(p.dir, p.ori, p.box, family=p.family, wrapped=wrapped, bc_type=bt)
            lat.setBoundaryDensity(pid, p.values)
    if case.alpha is not None:
        lat.setPorosity(case.alpha)
    return lat


def gpu_run(case, f_storage, nsteps, stats=True, initialize=False, wrapped=False):
    """Uploads a storage-typed state, runs nsteps on the GPU, returns (state, stats).
    initialize: one pass of the data processors first (MultiBlock3D::initialize);
    wrapped: planes of the wrapped factories (downloads return them completed)."""
    lat = gpu_lattice(case, wrapped=wrapped)
    try:
        lat.setPopulations(f_storage)
        if initialize:
            lat.initialize()
        lat.toggleInternalStatistics(stats)
        lat.collideAndStream(nsteps)
        dtype = np.uint16 if case.storage in ("f16", "bf16") else (
            np.float64 if case.storage == "f64" else np.float32)
        out = lat.getPopulations(dtype=dtype)
        st = lat.getStatistics() if (stats and nsteps > 0) else None
        return out, st
    finally:
        lat.close()
