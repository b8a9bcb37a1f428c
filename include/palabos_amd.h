/* palabos_amd — MI355X-native D3Q19 collide-and-stream engine, C ABI.
 *
 * Every entry point returns 0 on success and a negative code on failure; the
 * message of the last failure on the calling thread is available from
 * plbamd_last_error(). No C++ exception crosses this boundary. All pointers
 * are host pointers unless stated otherwise; sizes are in bytes or cells as
 * named. Calls on one lattice must be serialised by the caller (Palabos is
 * single-threaded per rank, multiBlock/multiBlockLattice3D.hh:318-360).
 *
 * Populations cross the boundary in Palabos's static cell layout: C order
 * [nx][ny][nz][19] of shifted populations f-bar = f - t_i, D3Q19 velocity order
 * of latticeBoltzmann/nearestNeighborLattices3D.hh:166-178
 * (BlockLatticeDataTransfer3D::send, atomicBlock/blockLattice3D.hh:650-690).
 */
#ifndef PALABOS_AMD_H
#define PALABOS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes. */
#define PLBAMD_OK 0
#define PLBAMD_EINVAL (-1)   /* bad argument                                  */
#define PLBAMD_ENOMEM (-2)   /* device allocation failed                      */
#define PLBAMD_EHIP (-3)     /* HIP runtime error                             */
#define PLBAMD_ESTATE (-4)   /* call not valid in the lattice's current state */

/* Storage precision of the populations on the device. F16/BF16 store f-bar in
 * 16 bits and compute in FP32; F64 and F32 compute in their own precision. */
#define PLBAMD_F64 0
#define PLBAMD_F32 1
#define PLBAMD_F16 2
#define PLBAMD_BF16 3

/* Bulk (background) dynamics of a lattice. */
#define PLBAMD_BGK 0   /* BGKdynamics, basicDynamics/isoThermalDynamics.hh:318-330             */
#define PLBAMD_RBGK 1  /* RegularizedBGKdynamics, basicDynamics/isoThermalDynamics.hh:1587-1600 */

/* Per-cell dynamics tags accepted by plbamd_set_tags. */
#define PLBAMD_TAG_BULK 0        /* background dynamics                                     */
#define PLBAMD_TAG_BOUNCE_BACK 1 /* BounceBack (full-way, node), core/dynamics.hh:783-790     */
#define PLBAMD_TAG_NO_DYNAMICS 2 /* NoDynamics, core/dynamics.hh:968-971                      */
#define PLBAMD_TAG_MOMENTUM_EXCHANGE 3 /* MomentumExchangeBounceBack, bounceBackModels.hh:98-135 */
#define PLBAMD_TAG_WALL_LINK 4   /* solid for link-wise half-way bounce-back: neighbouring
                                    fluid cells reflect their outgoing populations in the same
                                    step, as Palabos does at a non-periodic domain edge
                                    (blockLattice3D.hh:403-427); the cell itself is not updated */

/* Boundary-plane families and kinds (plbamd_add_plane). */
#define PLBAMD_ZOUHE 0        /* ZouHeVelocity/PressureDynamics, boundaryCondition/zouHeDynamics.hh:44-107 */
#define PLBAMD_REGULARIZED 1  /* RegularizedVelocity/DensityBoundaryDynamics,
                                 boundaryCondition/regularizedBoundaryDynamics.hh:94-106,166-178 */
#define PLBAMD_WRAPPED 0x100    /* or-ed into the family: the wrapped factories
                                   (createZouHeBoundaryCondition3D, zouHeBoundary3D.hh:338;
                                   createLocalBoundaryCondition3D, boundaryCondition3D.hh:1163),
                                   whose post-stream processor leaves completed populations in
                                   the lattice (wrappedLocalBoundaryProcessor3D.hh:41-62):
                                   downloads return the plane cells completed */
#define PLBAMD_VELOCITY 0
#define PLBAMD_PRESSURE 1

/* Boundary types (boundary::BcType, boundaryCondition/boundaryCondition.h:40), set
 * with plbamd_set_boundary_type. Velocity planes, edges and corners: NEUMANN and
 * OUTFLOW copy the velocity of the cell one step inside into the boundary after
 * every step (CopyVelocityFunctional3D, neumannCondition3D.hh:116-128), FREESLIP
 * its tangential components (the normal ones zeroed, :151-172), NORMAL_OUTFLOW
 * its normal components (:196-217); pressure planes: NEUMANN copies the density
 * (CopyDensityFunctional3D, :241-252). DIRICHLET and DENSITY keep the value set
 * (boundaryInstantiator3D.h:1042-1063, :1097-1102, :1139-1157, :1216-1232). */
#define PLBAMD_BC_DIRICHLET 0
#define PLBAMD_BC_NEUMANN 1
#define PLBAMD_BC_FREESLIP 2
#define PLBAMD_BC_DENSITY 3
#define PLBAMD_BC_OUTFLOW 4
#define PLBAMD_BC_NORMAL_OUTFLOW 5

/* Statistics layout returned by plbamd_get_statistics (FP64). */
#define PLBAMD_STAT_AVG_RHOBAR 0 /* average pre-collision rhoBar of cells taking statistics   */
#define PLBAMD_STAT_AVG_USQR 1   /* average pre-collision |u|^2                               */
#define PLBAMD_STAT_MAX_USQR 2   /* maximum pre-collision |u|^2                               */
#define PLBAMD_STAT_FORCE_X 3    /* momentum-exchange force on the obstacle                   */
#define PLBAMD_STAT_FORCE_Y 4
#define PLBAMD_STAT_FORCE_Z 5
#define PLBAMD_STAT_COUNT 6      /* number of cells that took statistics                      */
#define PLBAMD_STAT_POROUS_X 7   /* momentum removed by the Brinkman term (x, y, z)           */
#define PLBAMD_STAT_POROUS_Y 8
#define PLBAMD_STAT_POROUS_Z 9
#define PLBAMD_NUM_STATS 10

typedef struct {
    int32_t x0, x1, y0, y1, z0, z1; /* inclusive, like plb::Box3D (core/geometry3D.h) */
} plbamd_box;

typedef struct {
    int32_t nx, ny, nz;   /* lattice extent in cells                                   */
    int32_t storage;      /* PLBAMD_F64 | F32 | F16 | BF16                             */
    int32_t dynamics;     /* PLBAMD_BGK | PLBAMD_RBGK                                  */
    int32_t periodic[3];  /* 1 = periodic along x/y/z, 0 = same-step edge reflection   */
    double omega;         /* relaxation frequency                                      */
    double force[3];      /* uniform Guo body force (ForcedD3Q19, externalForceDynamics.hh:215-232);
                             all zero = plain dynamics                                */
} plbamd_lattice_desc;

/* Selects and initialises the HIP device for the calling process. */
int plbamd_init(int device);

/* Message of the last failure on this thread ("" if none). */
const char* plbamd_last_error(void);

/* Creates a lattice (one Palabos block, envelope included) and returns its handle.
 * Replaces CoProcessor3D<T>::addDomain (coProcessors/coProcessor3D.h:77) and the
 * BlockLattice3D(nx,ny,nz,Dynamics*) constructor (atomicBlock/blockLattice3D.h:111).
 * Populations start at rest (f-bar = 0). */
int plbamd_lattice_create(const plbamd_lattice_desc* desc, int* handle);

/* Creates a multi-block lattice: the domain desc->nx x ny x nz tiled by
 * `nblocks` uneven blocks (inclusive global boxes that must cover the domain
 * without overlap). Replaces the MultiBlockLattice3D(MultiBlockManagement3D,
 * ...) constructor over a SparseBlockStructure3D built with addBlock
 * (multiBlock/multiBlockLattice3D.h:75-84, sparseBlockStructure3D.h:55-57).
 * On one device the tiling needs no envelopes: the blocks are merged into the
 * fewest blocks below 2^32 cells each -- one block (a single-block lattice,
 * bit-identical results) whenever the domain has fewer cells; above that,
 * face-adjacent blocks with equal extents merge, z-neighbours first. All entry
 * points take global coordinates. At most 65535 blocks. */
int plbamd_multiblock_create(const plbamd_lattice_desc* desc, const plbamd_box* blocks, int nblocks,
                             int* handle);

/* plbamd_multiblock_create with flags. PLBAMD_KEEP_BLOCKS keeps the given
 * blocks as they are: every block is swept by the same single launch per step;
 * blocks exchange a two-cell envelope (Palabos: one cell, every step) after
 * every second step, which keeps the bulk bit-identical to a single-block
 * lattice; populations can be uploaded (and initialised) only after an even
 * number of steps (PLBAMD_ESTATE otherwise). This is the path lattices above
 * 2^32 cells per block take; the flag selects it for any lattice. No Palabos
 * counterpart (MultiBlockLattice3D keeps its blocks: one per MPI rank). */
#define PLBAMD_KEEP_BLOCKS 1
int plbamd_multiblock_create_ex(const plbamd_lattice_desc* desc, const plbamd_box* blocks, int nblocks, int flags,
                                int* handle);

/* Number of blocks the device runs the lattice as (1 for a single-block
 * lattice and for a coalesced tiling). Replaces
 * SparseBlockStructure3D::getNumBlocks (multiBlock/sparseBlockStructure3D.h:71)
 * for the blocks that exist on the device. */
int plbamd_get_block_count(int handle, int* count);

/* Releases a lattice. Replaces BlockLattice3D's destructor (blockLattice3D.hh:120-130). */
int plbamd_lattice_destroy(int handle);

/* Sets per-cell dynamics tags, C order [nx][ny][nz], PLBAMD_TAG_*.
 * Replaces defineDynamics(lattice, domain, Dynamics*) for the BounceBack,
 * NoDynamics and MomentumExchangeBounceBack prototypes
 * (dataProcessors/dataInitializerWrapper3D.hh:232-235); the momentum-exchange
 * fluid directions are derived as initializeMomentumExchange does
 * (boundaryCondition/bounceBackModels3D.hh:42-75). */
int plbamd_set_tags(int handle, const uint8_t* tags);

/* Adds a boundary plane (face of the box must be one cell thick along `dir`).
 * Replaces OnLatticeBoundaryCondition3D::add{Velocity,Pressure}Boundary{0,1,2}{N,P}
 * (boundaryCondition/boundaryCondition3D.h:40-200) of the factories
 * createDynamicsBasedZouHeBoundaryCondition3D (zouHeBoundary3D.hh:345) and
 * createDynamicsBasedLocalBoundaryCondition3D (boundaryCondition3D.hh:1170):
 * the completion runs inside the cell's collision. With PLBAMD_WRAPPED in
 * `family` it stands for the wrapped factories (same steps, completed
 * snapshots). Returns the plane id. */
int plbamd_add_plane(int handle, int family, int kind, int dir, int ori,
                     const plbamd_box* box, int* plane);

/* Adds an external velocity edge: the cells of `box` (one cell thick along the two
 * axes after `plane`, (plane+1)%3 and (plane+2)%3) take StoreVelocityDynamics over
 * the background dynamics and, after every stream, the finite-difference edge
 * closure OuterVelocityEdgeFunctional3D<plane, normal1, normal2>
 * (finiteDifferenceBoundaryProcessor3D.hh:130-217); normal1 / normal2 (-1 or +1)
 * are the outward normal along (plane+1)%3 / (plane+2)%3. Replaces
 * OnLatticeBoundaryCondition3D::addExternalVelocityEdge{0,1,2}{N,P}{N,P}
 * (boundaryCondition3D.h; boundaryInstantiator3D.h:1113-1153) of the Zou-He and
 * regularized factories, which install the same edge. Velocities (3 doubles per
 * cell, zero at first) through plbamd_set_plane_values with the returned id. Single-
 * block lattices without a body force; the stencil (the cells one step inside
 * along both normals, the neighbours along the edge) must lie in the lattice and be
 * background, plane, edge or corner cells (PLBAMD_EINVAL at the next step
 * otherwise). */
int plbamd_add_velocity_edge(int handle, int plane, int normal1, int normal2, const plbamd_box* box, int* id);

/* Adds an external velocity corner at (x, y, z) with outward normal (x_normal,
 * y_normal, z_normal), each -1 or +1: StoreVelocityDynamics and, after every
 * stream, OuterVelocityCornerFunctional3D (finiteDifferenceBoundaryProcessor3D.hh:
 * 221-313). Replaces OnLatticeBoundaryCondition3D::addExternalVelocityCorner{N,P}^3
 * (boundaryInstantiator3D.h:1197-1232). Velocity (3 doubles) through
 * plbamd_set_plane_values. Same limits as plbamd_add_velocity_edge. */
int plbamd_add_velocity_corner(int handle, int x_normal, int y_normal, int z_normal, int x, int y, int z, int* id);

/* Sets the boundary type of a plane, edge or corner (PLBAMD_BC_*; the bcType
 * argument of the add*Boundary / addExternalVelocity* methods,
 * boundaryCondition3D.h:40-200). Non-Dirichlet types: single-block lattices. */
int plbamd_set_boundary_type(int handle, int id, int bc_type);

/* One pass of the lattice's data processors (Neumann-type copies, edge and corner
 * closures) on the current state, as MultiBlock3D::initialize does before the
 * first step (multiBlock/multiBlock3D.cpp:330-332); every step runs them after its
 * stream. No effect on a lattice without edges, corners or non-Dirichlet types. */
int plbamd_execute_processors(int handle);

/* Sets the imposed values of a plane, edge or corner, x,y,z C order over its box:
 * velocity planes, edges and corners take 3 doubles (u) per cell, pressure planes
 * 1 double (rho) per cell.
 * Replaces setBoundaryVelocity / setBoundaryDensity
 * (dataProcessors/dataInitializerFunctional3D.hh:830-845, :878-888). */
int plbamd_set_plane_values(int handle, int plane, const double* values, size_t count);

/* Sets a Brinkman porosity field alpha in [0,1] per cell ([nx][ny][nz], FP64);
 * NULL removes it. The equilibrium velocity of a porous cell is (1-alpha) u.
 * No Palabos counterpart. */
int plbamd_set_porosity(int handle, const double* alpha);

/* Copies populations of `box` between host and device. `elem` is the host
 * element type (PLBAMD_F64 or PLBAMD_F32 values, or PLBAMD_F16/BF16 raw
 * 16-bit patterns when it equals the storage type); `bytes` must equal
 * box cells * 19 * sizeof(elem). Conversion to the storage type rounds to
 * nearest even. Replaces CoProcessor3D<T>::send / receive
 * (coProcessors/coProcessor3D.h:93,103) and Cell::operator[] access. */
int plbamd_upload(int handle, const plbamd_box* box, int elem, const void* data, size_t bytes);
int plbamd_download(int handle, const plbamd_box* box, int elem, void* data, size_t bytes);

/* Writes / reads the lattice as a Palabos checkpoint file, the format of
 * saveBinaryBlock / loadBinaryBlock (io/serializerIO_3D.h:48,58;
 * io/serializerIO_3D.cpp:37-66) for a MultiBlockLattice3D<T, D3Q19Descriptor>
 * (force_externals = 0) or <T, ForcedD3Q19Descriptor> (force_externals = 1)
 * with T = double (element PLBAMD_F64) or float (PLBAMD_F32): the base64 text
 * of the data size (8 bytes, or 4 with enforce_uint, as the reference's
 * enforceUint), then the base64 text of every cell of the bounding box in C
 * order (x, y, z), each its 19 f-bar (+ the 3 force components) as T; cells in
 * no block are zeros (IndexOrdering::forward, the reference's default). The
 * device serializes and codes the cells. Loading checks the size header
 * against this lattice, rounds to the storage type (nearest even), requires
 * every cell's force externals to equal the lattice's force, and refreshes the
 * multi-block envelopes; a multi-block lattice loads after an even number of
 * steps. Errors: PLBAMD_EINVAL with a message (file, size or content). */
int plbamd_save_binary_block(int handle, const char* file_name, int element, int force_externals, int enforce_uint);
int plbamd_load_binary_block(int handle, const char* file_name, int element, int force_externals, int enforce_uint);

/* Sets the populations of `box` to equilibrium for per-cell rho and u
 * (FP64, box cells and box cells x 3). Replaces initializeAtEquilibrium
 * (dataProcessors/dataInitializerFunctional3D.hh:925-944). */
int plbamd_initialize_at_equilibrium(int handle, const plbamd_box* box,
                                     const double* rho, const double* u);

/* Advances the lattice by nsteps collide-and-stream steps (blocking).
 * Replaces CoProcessor3D<T>::collideAndStream (coProcessors/coProcessor3D.h:115)
 * and BlockLattice3D::collideAndStream() (atomicBlock/blockLattice3D.hh:253-330). */
int plbamd_collide_and_stream(int handle, int nsteps);

/* As plbamd_collide_and_stream, and returns the elapsed device time of the
 * steps in milliseconds, measured with HIP events on the lattice's stream. */
int plbamd_collide_and_stream_timed(int handle, int nsteps, double* ms);

/* Turns the per-step statistics on or off (on by default). Replaces
 * MultiBlock3D::toggleInternalStatistics (multiBlock/multiBlock3D.cpp:455). */
int plbamd_toggle_statistics(int handle, int on);

/* Statistics of the last step, PLBAMD_NUM_STATS doubles (PLBAMD_STAT_*).
 * Replaces getStoredAverageDensity / getStoredAverageEnergy and the
 * momentum-exchange sums of BlockStatistics (core/blockStatistics.cpp:77-120). */
int plbamd_get_statistics(int handle, double* out);

/* Per-cell density and velocity of `box` (FP64; box cells and box cells x 3) as
 * Palabos's Dynamics::computeRhoBar / computeVelocity report them: bulk cells
 * rho = 1 + sum f-bar, u = j / rho (+ F/2 with a body force); boundary-plane
 * cells the Dirichlet density and velocity; solid nodes rho = 1, u = 0
 * (momentTemplates3D.h:375-432, boundaryDynamics.hh:643-835, dynamics.hh:824-1058). */
int plbamd_compute_macroscopic(int handle, const plbamd_box* box, double* rho, double* u);

/* FP64 sums over `box` of the same per-cell quantities: out[0] = sum rhoBar,
 * out[1] = sum |u|^2, out[2..4] = sum u. computeAverageDensity is
 * 1 + out[0] / cells and computeAverageEnergy out[1] / (2 cells)
 * (dataProcessors/dataAnalysisWrapper3D.hh:1226-1257). */
int plbamd_reduce_moments(int handle, const plbamd_box* box, double* out);

/* plbamd_reduce_moments without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * captured into a HIP graph, tight serving loops, harnesses that forbid
 * hipMalloc). The per-workgroup partial records go to a fixed device array of
 * the library (8 MiB per GPU, held until the process ends) instead of memory
 * allocated per call; the call makes no hipMalloc / hipFree. Same kernels and
 * the same fixed-order reduction, so `out` equals plbamd_reduce_moments's bit
 * for bit. Limits: at most 131072 partial records (every single-block lattice
 * fits; a multi-block box needing more is refused with PLBAMD_EINVAL), and
 * PLBAMD_ESTATE when the lattice's tags, planes or porosity changed since its
 * last step or plbamd_reduce_moments (their device set-up allocates). Every
 * call on a GPU shares that GPU's array; the library serialises the calls of
 * one process, and each process has its own array. */
int plbamd_reduce_moments_noalloc(int handle, const plbamd_box* box, double* out);

/* Density parameter of the solid dynamics of tag PLBAMD_TAG_BOUNCE_BACK,
 * PLBAMD_TAG_NO_DYNAMICS or PLBAMD_TAG_MOMENTUM_EXCHANGE (the rho of
 * BounceBack(rho), NoDynamics(rho), MomentumExchangeBounceBack(forceIds, rho);
 * core/dynamics.h:574,687, bounceBackModels.h:62): what the analysis functions
 * report for those cells (computeDensity = rho; computeRhoBar = rho - 1, 0 for
 * NoDynamics). One value per kind; 1 until set. The collision does not use it. */
int plbamd_set_solid_density(int handle, int tag, double rho);

/* Analysis fields of `box` as Palabos's data-analysis functions compute them
 * (dataProcessors/dataAnalysisWrapper3D.h:695-777; per cell the cell's
 * dynamics decides, e.g. BounceBack reports rho = 1 and u = 0, a velocity
 * plane its imposed velocity): out[cell][7] (box cells in C order) =
 * computeDensity, computeRhoBar, computeVelocity (x, y, z),
 * computeKineticEnergy (|u|^2 / 2), computeVelocityNorm, each computed in the
 * lattice's arithmetic type (FP64, or FP32 for FP32 and 16-bit storage) and
 * widened exactly to double. */
int plbamd_compute_fields(int handle, const plbamd_box* box, double* out);

/* Objective of plbamd_adjoint_gradient:
 *   J = 1/(T - t0) sum_{t=t0+1..T} [ weight_pressure_drop * cs2 * (<rho>_{x=inlet_x} - <rho>_{x=outlet_x})
 *                                  + weight_drag * omega * sum_cells alpha * j_x
 *                                  + weight_flux * <j_x> ]
 * evaluated on the states after each step, t0 = first_step (plane means over
 * all y, z; the drag term is the momentum the Brinkman porosity removes in the
 * next collision, x component; <j_x> the mean x momentum over all cells, whose
 * time average times nu / g is the Darcy permeability of a force-driven
 * porous medium; FP64 accumulation). */
typedef struct {
    int32_t inlet_x, outlet_x;
    double weight_pressure_drop;
    double weight_drag;
    double weight_flux;
    int32_t first_step;
} plbamd_objective;

/* Runs nsteps steps from the current state (the lattice advances) and returns J
 * and its gradient with respect to the porosity of every cell (grad_alpha,
 * [nx][ny][nz], may be NULL) and to the scale s of every plane's imposed
 * velocity, u -> s u at s = 1 (grad_plane_scale, one per plane in creation
 * order; 0 for pressure planes; may be NULL). Hand-written discrete adjoint of
 * the collide-and-stream (BGK or regularized BGK, Guo body force, Brinkman
 * porosity) with checkpoints every `checkpoint_every` steps (even) and
 * recomputation of the segments not recorded in the forward sweep
 * (plbamd_set_adjoint_record_bytes); timing_ms (may be NULL) receives the
 * forward and the backward device time. FP64 or 16-bit storage (the 16-bit forward is differentiated
 * through its FP32 arithmetic, the storage rounding taken as the identity;
 * every block even along z); of the per-cell tags only link-wise walls
 * (reflected as the forward reflects them; solid cells count zero in J, and
 * not together with planes); planes only on single-block BGK lattices without a body force, with
 * zero porosity on them; started after an even number of steps. Multi-block
 * lattices refresh the first envelope layer of the adjoint state after every
 * reverse step.
 * No Palabos counterpart. */
int plbamd_adjoint_gradient(int handle, int nsteps, int checkpoint_every, const plbamd_objective* objective,
                            double* J, double* grad_alpha, double* grad_plane_scale, double* timing_ms);

/* Device memory plbamd_adjoint_gradient may use for records made during its
 * forward sweep: the moments (and plane populations) of every step of the
 * last segments, which then need no recompute in the reverse sweep (the same
 * kernels on the same states, so the same records and the same gradient).
 * bytes >= 0: at most that many; -1 (the default): the free device memory at
 * the call beyond a reserve of max(4 GiB, 1/32 of the device); 0: every
 * segment is recomputed. No Palabos counterpart. */
int plbamd_set_adjoint_record_bytes(int handle, int64_t bytes);

/* Segments the last plbamd_adjoint_gradient call recorded in its forward
 * sweep, and the device bytes of those records. No Palabos counterpart. */
int plbamd_get_adjoint_record_info(int handle, int* kept_segments, int64_t* bytes);

/* Number of completed steps. */
int plbamd_get_time(int handle, int64_t* t);

/* External moments: the lattice's rhoBar / j field (the ScalarField3D rhoBar and
 * TensorField3D<T,3> j that Palabos's processors share), kept on the device in
 * the storage's arithmetic type (FP64 for FP64 storage, FP32 otherwise),
 * allocated on first use. Supported on single-block lattices of bulk BGK /
 * RBGK cells with BounceBack and NoDynamics cells; other lattices are refused
 * with PLBAMD_EINVAL. Host arrays: C order over the box, j with 3 values per
 * cell.
 * plbamd_compute_rhobar_j replaces applying BoxRhoBarJfunctional3D on `box`
 * (dataProcessors/dataAnalysisFunctional3D.hh:528-561): every cell's
 * computeRhoBarJ (bulk: get_rhoBar_j's partial sums; BounceBack: its density
 * - 1 and 0; NoDynamics: 0 and 0). plbamd_get_rhobar_j / plbamd_set_rhobar_j
 * read and write the field (ScalarField3D::get / TensorField3D::get); values
 * set are rounded to the arithmetic type. */
int plbamd_compute_rhobar_j(int handle, const plbamd_box* box);
int plbamd_get_rhobar_j(int handle, const plbamd_box* box, double* rho_bar, double* j);
int plbamd_set_rhobar_j(int handle, const plbamd_box* box, const double* rho_bar, const double* j);

/* nsteps steps of ExternalRhoJcollideAndStream3D over the whole lattice
 * (basicDynamics/dynamicsProcessor3D.hh:41-209) with the field as it is, on
 * lattices periodic along every axis (at a non-periodic edge the reference
 * streams through the block's envelope layer, dynamicsProcessor3D.hh:118-190,
 * so populations enter from outside the domain; such lattices are refused with
 * PLBAMD_EINVAL; BounceBack cells make the walls): every
 * cell collides with the field's moments (collideExternal: BGK or regularized
 * BGK with the given rhoBar and j, isoThermalDynamics.hh:332-338 and
 * :1604-1613; BounceBack swaps; NoDynamics nothing), then streams. No statistics are taken
 * (as collideExternal); the step counter advances. PLBAMD_ESTATE when the field
 * was never computed or set. */
int plbamd_external_collide_and_stream(int handle, int nsteps);

/* Streaming-bandwidth probe: the rate this GPU sustains for the memory access
 * patterns of the collide-and-stream kernels, with their nontemporal loads and
 * stores and no arithmetic; the practical ceiling their HBM fractions are read
 * against. PLBAMD_PROBE_COPY copies bytes/2 into a second buffer in 16-byte
 * elements; PLBAMD_PROBE_SLOTS reads and writes back in place 19 slot arrays
 * of bytes/38 each (an even step's data movement) in 8-byte elements, one per
 * thread (one FP64 cell); _SLOTS_4B and _SLOTS_16B in 4- and 16-byte ones
 * (one FP32 cell, two FP64 cells per thread). Allocates and
 * frees its buffers; one untimed launch, then `repeats` timed ones (HIP events
 * on its own stream). Out: the shortest launch in ms and the bytes each launch
 * moves (read + written). No Palabos counterpart: a measurement helper. */
#define PLBAMD_PROBE_COPY 0
#define PLBAMD_PROBE_SLOTS 1
#define PLBAMD_PROBE_SLOTS_4B 2
#define PLBAMD_PROBE_SLOTS_16B 3
int plbamd_bandwidth_probe(int pattern, uint64_t bytes, int repeats, double* ms_best, double* bytes_moved);

#ifdef __cplusplus
}
#endif

#endif /* PALABOS_AMD_H */
