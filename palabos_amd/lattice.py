"""Python host mirror of Palabos's lattice API for the D3Q19 hot path.

Names and argument meaning follow the reference's Python binding (pypal, as
used by jlabos/src/testing/test_cylinder.py:78-130: Block3D, BGK,
initializeAtEquilibrium, collideAndStream, averageEnergy, averageDensity) and
the C++ API (BlockLattice3D::collideAndStream, defineDynamics,
setBoundaryVelocity, setBoundaryDensity, getStoredAverageDensity/Energy).
Everything runs through the native library (include/palabos_amd.h); there is
no CPU path.

Boxes are inclusive (x0, x1, y0, y1, z0, z1) tuples, as plb::Box3D.
Populations are numpy arrays [x][y][z][19] of shifted f-bar in Palabos's
D3Q19 order.
"""
import ctypes
import os

import numpy as np

from . import native
from .native import Box, check, lib

D3Q19 = "D3Q19"



def _file_elem(T):
    codes = {"f64": native.F64, "double": native.F64, "f32": native.F32, "float": native.F32}
    if T not in codes:
        raise ValueError("checkpoint element type must be f64 or f32")
    return codes[T]

class BGK:
    """BGKdynamics (basicDynamics/isoThermalDynamics.hh:318-330)."""
    code = native.BGK

    def __init__(self, omega):
        self.omega = float(omega)


class RegularizedBGK:
    """RegularizedBGKdynamics (basicDynamics/isoThermalDynamics.hh:1587-1600)."""
    code = native.RBGK

    def __init__(self, omega):
        self.omega = float(omega)


class BounceBack:
    """BounceBack node (core/dynamics.hh:783-790)."""
    tag = native.TAG_BOUNCE_BACK


class NoDynamics:
    """NoDynamics (core/dynamics.hh:968-971)."""
    tag = native.TAG_NO_DYNAMICS


class MomentumExchangeBounceBack:
    """MomentumExchangeBounceBack (boundaryCondition/bounceBackModels.hh:98-135)."""
    tag = native.TAG_MOMENTUM_EXCHANGE


class WallLink:
    """Solid of link-wise half-way bounce-back (same-step reflection of the
    neighbours' outgoing populations, as at a non-periodic domain edge)."""
    tag = native.TAG_WALL_LINK


class Bulk:
    tag = native.TAG_BULK


def _box(b, shape):
    if b is None:
        return Box(0, shape[0] - 1, 0, shape[1] - 1, 0, shape[2] - 1)
    return Box.of(b)


class Block3D:
    """A lattice: one block (a Palabos BlockLattice3D / single-block
    MultiBlockLattice3D) or, with `blocks`, a multi-block lattice tiled by uneven
    inclusive boxes (MultiBlockLattice3D over a SparseBlockStructure3D). The
    device runs a tiling as one block while it has fewer than 2^32 cells
    (bit-identical results); keep_blocks=True keeps the given blocks, with
    their envelope exchange."""

    def __init__(self, nx, ny, nz, descriptor=D3Q19, dynamics=None, storage="f64",
                 periodic=False, force=None, device=0, blocks=None, keep_blocks=False):
        if descriptor != D3Q19:
            raise ValueError("only the D3Q19 descriptor is supported")
        if dynamics is None:
            raise ValueError("a bulk dynamics (BGK or RegularizedBGK) is required")
        native.init(device)
        if isinstance(periodic, bool):
            periodic = (periodic,) * 3
        d = native.LatticeDesc()
        d.nx, d.ny, d.nz = int(nx), int(ny), int(nz)
        d.storage = native.STORAGE_CODES[storage]
        d.dynamics = dynamics.code
        for a in range(3):
            d.periodic[a] = int(bool(periodic[a]))
        d.omega = dynamics.omega
        if force is not None:
            for a in range(3):
                d.force[a] = float(force[a])
        h = ctypes.c_int(0)
        if blocks is None:
            check(lib().plbamd_lattice_create(ctypes.byref(d), ctypes.byref(h)))
        else:
            arr = (Box * len(blocks))(*[Box.of(b) for b in blocks])
            flags = native.KEEP_BLOCKS if keep_blocks else 0
            check(lib().plbamd_multiblock_create_ex(ctypes.byref(d), arr, len(blocks), flags, ctypes.byref(h)))
        self.blocks = None if blocks is None else [tuple(b) for b in blocks]
        self.handle = h.value
        self.shape = (int(nx), int(ny), int(nz))
        self.storage = storage
        self.dynamics = dynamics
        self.periodic = tuple(bool(p) for p in periodic)
        self._tags = None
        self._planes = []
        self._plane_values = {}   # velocity boundaries: the values last set from here, by id

    def blockCount(self):
        """Blocks the device runs this lattice as (1 for a coalesced tiling)."""
        n = ctypes.c_int(0)
        check(lib().plbamd_get_block_count(self.handle, ctypes.byref(n)))
        return n.value

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "handle", None) is not None:
            check(lib().plbamd_lattice_destroy(self.handle))
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nx(self):
        return self.shape[0]

    @property
    def ny(self):
        return self.shape[1]

    @property
    def nz(self):
        return self.shape[2]

    def getBoundingBox(self):
        return (0, self.nx - 1, 0, self.ny - 1, 0, self.nz - 1)

    # -- geometry / dynamics ------------------------------------------------
    def setTags(self, tags):
        tags = np.ascontiguousarray(tags, dtype=np.uint8)
        if tags.shape != self.shape:
            raise ValueError("tags must have the lattice shape")
        self._tags = tags.copy()
        check(lib().plbamd_set_tags(self.handle, tags.ctypes.data))

    def defineDynamics(self, where, dynamics):
        """defineDynamics(lattice, box|mask, dynamics) (dataInitializerWrapper3D.hh:232-240)."""
        if self._tags is None:
            self._tags = np.zeros(self.shape, dtype=np.uint8)
        if isinstance(where, np.ndarray):
            self._tags[where.astype(bool)] = dynamics.tag
        else:
            x0, x1, y0, y1, z0, z1 = where
            self._tags[x0:x1 + 1, y0:y1 + 1, z0:z1 + 1] = dynamics.tag
        check(lib().plbamd_set_tags(self.handle, self._tags.ctypes.data))

    def _add_plane(self, family, kind, direction, orientation, box, wrapped=False, bc_type="dirichlet"):
        fam = {"zouhe": native.ZOUHE, "regularized": native.REGULARIZED}[family]
        if wrapped:
            fam |= native.WRAPPED
        pid = ctypes.c_int(-1)
        b = Box.of(box)
        check(lib().plbamd_add_plane(self.handle, fam, kind, int(direction), int(orientation),
                                     ctypes.byref(b), ctypes.byref(pid)))
        self._planes.append((kind, b))
        self.setBoundaryType(pid.value, bc_type)
        return pid.value

    def addVelocityBoundary(self, direction, orientation, box, family="zouhe", wrapped=False, bc_type="dirichlet"):
        """add{Velocity}Boundary{dir}{N|P}(box, lattice, bcType) of the Zou-He or
        regularized boundary condition (boundaryCondition3D.h:40-200): the
        dynamics-based factory, or with wrapped=True the wrapped one
        (createZouHeBoundaryCondition3D / createLocalBoundaryCondition3D). Returns
        a boundary id."""
        return self._add_plane(family, native.VELOCITY, direction, orientation, box, wrapped, bc_type)

    def addPressureBoundary(self, direction, orientation, box, family="zouhe", wrapped=False, bc_type="dirichlet"):
        return self._add_plane(family, native.PRESSURE, direction, orientation, box, wrapped, bc_type)

    def addExternalVelocityEdge(self, plane, normal1, normal2, box, bc_type="dirichlet"):
        """addExternalVelocityEdge{plane}{normal1}{normal2}(box, lattice, bcType)
        (boundaryInstantiator3D.h:1113-1153): StoreVelocityDynamics and the
        finite-difference edge closure; normal1 / normal2 (-1 | +1) are the outward
        normal along (plane+1)%3 / (plane+2)%3. Returns a boundary id."""
        pid = ctypes.c_int(-1)
        b = Box.of(box)
        check(lib().plbamd_add_velocity_edge(self.handle, int(plane), int(normal1), int(normal2), ctypes.byref(b),
                                             ctypes.byref(pid)))
        self._planes.append((native.VELOCITY, b))
        self.setBoundaryType(pid.value, bc_type)
        return pid.value

    def addExternalVelocityCorner(self, x_normal, y_normal, z_normal, x, y, z, bc_type="dirichlet"):
        """addExternalVelocityCorner{xN}{yN}{zN}(x, y, z, lattice, bcType)
        (boundaryInstantiator3D.h:1197-1232). Returns a boundary id."""
        pid = ctypes.c_int(-1)
        check(lib().plbamd_add_velocity_corner(self.handle, int(x_normal), int(y_normal), int(z_normal),
                                               int(x), int(y), int(z), ctypes.byref(pid)))
        self._planes.append((native.VELOCITY, Box(int(x), int(x), int(y), int(y), int(z), int(z))))
        self.setBoundaryType(pid.value, bc_type)
        return pid.value

    def setVelocityConditionOnBlockBoundaries(self, family="regularized", wrapped=True, bc_type="dirichlet"):
        """OnLatticeBoundaryCondition3D::setVelocityConditionOnBlockBoundaries(lattice,
        bcType) (boundaryCondition3D.hh:300-400): velocity conditions on the six face
        interiors, the twelve edges and the eight corners of the whole lattice, in
        that order. Returns the 26 boundary ids."""
        n = self.shape
        ids = []
        for d in range(3):
            for ori in (-1, 1):
                lo = [1, 1, 1]
                hi = [n[0] - 2, n[1] - 2, n[2] - 2]
                lo[d] = hi[d] = 0 if ori < 0 else n[d] - 1
                ids.append(self.addVelocityBoundary(d, ori, (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]), family,
                                                    wrapped, bc_type))
        for plane in range(3):
            d1, d2 = (plane + 1) % 3, (plane + 2) % 3
            for n1 in (-1, 1):
                for n2 in (-1, 1):
                    lo = [0, 0, 0]
                    hi = [0, 0, 0]
                    lo[plane], hi[plane] = 1, n[plane] - 2
                    lo[d1] = hi[d1] = 0 if n1 < 0 else n[d1] - 1
                    lo[d2] = hi[d2] = 0 if n2 < 0 else n[d2] - 1
                    ids.append(self.addExternalVelocityEdge(plane, n1, n2, (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]),
                                                            bc_type))
        for xn in (-1, 1):
            for yn in (-1, 1):
                for zn in (-1, 1):
                    ids.append(self.addExternalVelocityCorner(xn, yn, zn, 0 if xn < 0 else n[0] - 1,
                                                              0 if yn < 0 else n[1] - 1, 0 if zn < 0 else n[2] - 1,
                                                              bc_type))
        return ids

    def setBoundaryType(self, boundary, bc_type):
        """bcType of a plane, edge or corner: 'dirichlet', 'neumann', 'freeslip',
        'density', 'outflow' or 'normalOutflow' (boundary::BcType)."""
        check(lib().plbamd_set_boundary_type(self.handle, int(boundary), native.BC_TYPES[bc_type]))

    def initialize(self):
        """MultiBlock3D::initialize: one pass of the data processors (copies of
        Neumann-type boundaries, edge and corner closures)."""
        check(lib().plbamd_execute_processors(self.handle))

    def setBoundaryVelocityIn(self, box, u):
        """setBoundaryVelocity(lattice, box, u) (dataInitializerWrapper3D.h): u on every
        velocity plane, edge and corner cell inside the box."""
        x0, x1, y0, y1, z0, z1 = box
        u = np.asarray(u, dtype=np.float64)
        for pid, (kind, b) in enumerate(self._planes):
            if kind != native.VELOCITY:
                continue
            ix0, ix1 = max(x0, b.x0), min(x1, b.x1)
            iy0, iy1 = max(y0, b.y0), min(y1, b.y1)
            iz0, iz1 = max(z0, b.z0), min(z1, b.z1)
            if ix0 > ix1 or iy0 > iy1 or iz0 > iz1:
                continue
            vals = self._plane_values.get(pid)
            if vals is None:
                vals = np.zeros(b.shape() + (3,))
            vals[ix0 - b.x0:ix1 - b.x0 + 1, iy0 - b.y0:iy1 - b.y0 + 1, iz0 - b.z0:iz1 - b.z0 + 1] = u
            self._plane_values[pid] = vals
            flat = np.ascontiguousarray(vals.reshape(-1, 3))
            check(lib().plbamd_set_plane_values(self.handle, pid, flat.ctypes.data, flat.size))

    def setBoundaryVelocity(self, plane, u):
        """setBoundaryVelocity of one velocity plane, edge or corner: a 3-vector or an
        array [box cells, 3] in box C order."""
        kind, b = self._planes[plane]
        if kind != native.VELOCITY:
            raise ValueError("not a velocity boundary")
        n = b.cells()
        u = np.asarray(u, dtype=np.float64)
        vals = np.broadcast_to(u, (n, 3)) if u.shape == (3,) else u.reshape(n, 3)
        vals = np.ascontiguousarray(vals)
        self._plane_values[plane] = vals.reshape(b.shape() + (3,)).copy()
        check(lib().plbamd_set_plane_values(self.handle, plane, vals.ctypes.data, vals.size))

    def setBoundaryDensity(self, plane, rho):
        kind, b = self._planes[plane]
        if kind != native.PRESSURE:
            raise ValueError("not a pressure plane")
        n = b.cells()
        vals = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, dtype=np.float64), (n,)).reshape(n))
        check(lib().plbamd_set_plane_values(self.handle, plane, vals.ctypes.data, vals.size))

    def setPorosity(self, alpha):
        """Brinkman porosity field alpha in [0,1] (None removes it)."""
        if alpha is None:
            check(lib().plbamd_set_porosity(self.handle, None))
            return
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError("alpha must have the lattice shape")
        check(lib().plbamd_set_porosity(self.handle, a.ctypes.data))

    # -- data ---------------------------------------------------------------
    def initializeAtEquilibrium(self, rho, u, box=None):
        b = _box(box, self.shape)
        n = b.cells()
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(rho, dtype=np.float64), b.shape()).reshape(n))
        uu = np.asarray(u, dtype=np.float64)
        uu = np.ascontiguousarray(np.broadcast_to(uu, b.shape() + (3,)).reshape(n, 3))
        check(lib().plbamd_initialize_at_equilibrium(self.handle, ctypes.byref(b), r.ctypes.data,
                                                      uu.ctypes.data))

    def setPopulations(self, f, box=None):
        """Uploads populations (float64/float32 values, or raw uint16 bits of the storage type)."""
        b = _box(box, self.shape)
        f = np.ascontiguousarray(f)
        if f.shape != b.shape() + (19,):
            raise ValueError(f"populations must have shape {b.shape() + (19,)}")
        elem = {np.dtype(np.float64): native.F64, np.dtype(np.float32): native.F32}.get(f.dtype)
        if elem is None:
            if f.dtype != np.uint16 or self.storage not in ("f16", "bf16"):
                raise ValueError("unsupported population dtype")
            elem = native.STORAGE_CODES[self.storage]
        check(lib().plbamd_upload(self.handle, ctypes.byref(b), elem, f.ctypes.data, f.nbytes))

    def getPopulations(self, box=None, dtype=np.float64):
        """Downloads populations as float64/float32 values, or as raw storage bits (dtype=np.uint16)."""
        b = _box(box, self.shape)
        dtype = np.dtype(dtype)
        if dtype == np.float64:
            elem = native.F64
        elif dtype == np.float32:
            elem = native.F32
        elif dtype == np.uint16 and self.storage in ("f16", "bf16"):
            elem = native.STORAGE_CODES[self.storage]
        else:
            raise ValueError("unsupported population dtype")
        out = np.empty(b.shape() + (19,), dtype=dtype)
        check(lib().plbamd_download(self.handle, ctypes.byref(b), elem, out.ctypes.data, out.nbytes))
        return out

    # -- checkpoint files -----------------------------------------------------
    def saveBinaryBlock(self, fName, T="f64", forceExternals=False, enforceUint=False):
        """Writes the lattice as a Palabos checkpoint file (saveBinaryBlock,
        io/serializerIO_3D.cpp:37-49): the file a MultiBlockLattice3D<T, D3Q19>
        (or <T, ForcedD3Q19> with forceExternals) of the same populations writes;
        T is "f64" (double) or "f32" (float)."""
        check(lib().plbamd_save_binary_block(self.handle, os.fsencode(fName), _file_elem(T), int(bool(forceExternals)),
                                             int(bool(enforceUint))))

    def loadBinaryBlock(self, fName, T="f64", forceExternals=False, enforceUint=False):
        """Reads a Palabos checkpoint file (loadBinaryBlock,
        io/serializerIO_3D.cpp:51-66) into the lattice."""
        check(lib().plbamd_load_binary_block(self.handle, os.fsencode(fName), _file_elem(T), int(bool(forceExternals)),
                                             int(bool(enforceUint))))

    # -- stepping -----------------------------------------------------------
    def collideAndStream(self, nsteps=1):
        check(lib().plbamd_collide_and_stream(self.handle, int(nsteps)))

    def collideAndStreamTimed(self, nsteps):
        """Runs nsteps steps and returns the device time in ms (HIP events on the lattice stream)."""
        ms = ctypes.c_double(0.0)
        check(lib().plbamd_collide_and_stream_timed(self.handle, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def adjointGradient(self, nsteps, checkpoint_every, inlet_x=0, outlet_x=0, weight_pressure_drop=1.0,
                        weight_drag=1.0, weight_flux=0.0, first_step=0):
        """Runs nsteps (the lattice advances) and returns the objective
        J = mean over steps t0+1..T of [w_dp cs2 (<rho>_inlet_x - <rho>_outlet_x)
        + w_drag omega sum alpha j_x + w_flux <j_x>] (t0 = first_step) with its
        gradient with respect to the porosity field and to the scale of each
        plane's imposed velocity (hand-written adjoint, checkpoints every
        `checkpoint_every` steps). Returns (J, grad_alpha [nx,ny,nz], grad_plane_scale
        [planes], (forward_ms, backward_ms))."""
        obj = native.Objective(int(inlet_x), int(outlet_x), float(weight_pressure_drop), float(weight_drag),
                               float(weight_flux), int(first_step))
        J = ctypes.c_double(0.0)
        ga = np.zeros(self.shape, dtype=np.float64)
        gs = np.zeros(max(1, len(self._planes)), dtype=np.float64)
        tm = np.zeros(2, dtype=np.float64)
        check(lib().plbamd_adjoint_gradient(self.handle, int(nsteps), int(checkpoint_every), ctypes.byref(obj),
                                            ctypes.byref(J), ga.ctypes.data, gs.ctypes.data, tm.ctypes.data))
        return J.value, ga, gs[:len(self._planes)], (tm[0], tm[1])

    def setAdjointRecordBytes(self, nbytes):
        """Device memory adjointGradient may use for the records of the last
        segments made during its forward sweep (those segments are then not
        recomputed; the gradient is the same): -1 (default) the free memory
        beyond a reserve, 0 none, otherwise at most nbytes."""
        check(lib().plbamd_set_adjoint_record_bytes(self.handle, int(nbytes)))

    def adjointRecordInfo(self):
        """(segments recorded in the last adjointGradient's forward sweep, their bytes)."""
        n, b = ctypes.c_int(0), ctypes.c_int64(0)
        check(lib().plbamd_get_adjoint_record_info(self.handle, ctypes.byref(n), ctypes.byref(b)))
        return n.value, b.value

    def toggleInternalStatistics(self, on):
        check(lib().plbamd_toggle_statistics(self.handle, int(bool(on))))

    @property
    def time(self):
        t = ctypes.c_int64(0)
        check(lib().plbamd_get_time(self.handle, ctypes.byref(t)))
        return t.value

    # -- analysis -----------------------------------------------------------
    def getStatistics(self):
        s = np.zeros(native.NUM_STATS, dtype=np.float64)
        check(lib().plbamd_get_statistics(self.handle, s.ctypes.data))
        return {"avg_rhobar": s[0], "avg_usqr": s[1], "max_usqr": s[2], "force": s[3:6].copy(),
                "count": int(s[6]), "porous": s[7:10].copy()}

    def getStoredAverageDensity(self):
        """multiBlockLattice3D.hh:531-535."""
        return 1.0 + self.getStatistics()["avg_rhobar"]

    def getStoredAverageEnergy(self):
        """multiBlockLattice3D.hh:538-541."""
        return 0.5 * self.getStatistics()["avg_usqr"]

    def macroscopic(self, box=None):
        b = _box(box, self.shape)
        n = b.cells()
        rho = np.empty(n, dtype=np.float64)
        u = np.empty((n, 3), dtype=np.float64)
        check(lib().plbamd_compute_macroscopic(self.handle, ctypes.byref(b), rho.ctypes.data,
                                                u.ctypes.data))
        return rho.reshape(b.shape()), u.reshape(b.shape() + (3,))

    def density(self, box=None):
        return self.macroscopic(box)[0]

    def setSolidDensity(self, dynamics, rho):
        """Density parameter of a solid dynamics (BounceBack, NoDynamics or
        MomentumExchangeBounceBack, class or tag): what the analysis functions
        report for its cells (1 until set; Palabos's constructors default to 0)."""
        tag = getattr(dynamics, "tag", dynamics)
        check(lib().plbamd_set_solid_density(self.handle, int(tag), float(rho)))

    # -- analysis fields (dataAnalysisWrapper3D.h:695-777) --------------------
    def analysisFields(self, box=None):
        """[nx, ny, nz, 7] of the box: computeDensity, computeRhoBar,
        computeVelocity (x, y, z), computeKineticEnergy, computeVelocityNorm,
        in the lattice's arithmetic type widened to float64."""
        b = _box(box, self.shape)
        out = np.empty(b.shape() + (7,), dtype=np.float64)
        check(lib().plbamd_compute_fields(self.handle, ctypes.byref(b), out.ctypes.data))
        return out

    def computeDensity(self, box=None):
        return self.analysisFields(box)[..., 0]

    def computeRhoBar(self, box=None):
        return self.analysisFields(box)[..., 1]

    def computeVelocity(self, box=None):
        return self.analysisFields(box)[..., 2:5]

    def computeVelocityComponent(self, box=None, iComponent=0):
        return self.analysisFields(box)[..., 2 + int(iComponent)]

    def computeKineticEnergy(self, box=None):
        return self.analysisFields(box)[..., 5]

    def computeVelocityNorm(self, box=None):
        return self.analysisFields(box)[..., 6]

    def velocity(self, box=None):
        return self.macroscopic(box)[1]

    # -- external moments (dynamicsProcessor3D.hh:41-209, dataAnalysisFunctional3D.hh:528-561) --
    def computeRhoBarJ(self, box=None):
        """BoxRhoBarJfunctional3D on the box: the lattice's rhoBar / j field
        (on the device) takes every cell's computeRhoBarJ."""
        b = _box(box, self.shape)
        check(lib().plbamd_compute_rhobar_j(self.handle, ctypes.byref(b)))

    def getRhoBarJ(self, box=None):
        """The rhoBar / j field on the box: ([nx, ny, nz], [nx, ny, nz, 3]) float64."""
        b = _box(box, self.shape)
        rho_bar = np.empty(b.shape(), dtype=np.float64)
        j = np.empty(b.shape() + (3,), dtype=np.float64)
        check(lib().plbamd_get_rhobar_j(self.handle, ctypes.byref(b), rho_bar.ctypes.data, j.ctypes.data))
        return rho_bar, j

    def setRhoBarJ(self, rho_bar, j, box=None):
        """Sets the rhoBar / j field on the box (rounded to the arithmetic type)."""
        b = _box(box, self.shape)
        rho_bar = np.ascontiguousarray(rho_bar, dtype=np.float64).reshape(b.shape())
        j = np.ascontiguousarray(j, dtype=np.float64).reshape(b.shape() + (3,))
        check(lib().plbamd_set_rhobar_j(self.handle, ctypes.byref(b), rho_bar.ctypes.data, j.ctypes.data))

    def externalCollideAndStream(self, nsteps=1):
        """ExternalRhoJcollideAndStream3D: nsteps steps colliding with the
        rhoBar / j field as it is (no statistics)."""
        check(lib().plbamd_external_collide_and_stream(self.handle, int(nsteps)))

    def reduceMoments(self, box=None):
        b = _box(box, self.shape)
        out = np.zeros(5, dtype=np.float64)
        check(lib().plbamd_reduce_moments(self.handle, ctypes.byref(b), out.ctypes.data))
        return {"sum_rhobar": out[0], "sum_usqr": out[1], "sum_u": out[2:5].copy(), "cells": b.cells()}

    def averageDensity(self, box=None):
        """computeAverageDensity (dataAnalysisWrapper3D.hh:1226-1230)."""
        r = self.reduceMoments(box)
        return 1.0 + r["sum_rhobar"] / r["cells"]

    def averageEnergy(self, box=None):
        """computeAverageEnergy (dataAnalysisWrapper3D.hh:1252-1257): mean of |u|^2 / 2."""
        r = self.reduceMoments(box)
        return 0.5 * r["sum_usqr"] / r["cells"]
