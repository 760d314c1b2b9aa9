/*
 * mrmd_hip.h - C ABI of the MI355X (gfx950) implementation of MRMD's per-timestep hot path.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and returns an int
 * status (MRMD_OK == 0, negative on error; the message is in mrmd_ctx_last_error()).
 * No exception crosses the ABI.  Each context owns one HIP stream; calls are asynchronous
 * on it unless they return a host scalar (energies, ghost counts, displacements), in which
 * case they synchronise - the same fence points as the reference, whose every action ends
 * in Kokkos::fence().
 *
 * Reference citations (XzzX/mrmd v0.2.0, paths relative to its source tree) name the
 * interface each entry point replaces.
 */
#ifndef MRMD_HIP_H
#define MRMD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------- */
#define MRMD_OK 0
#define MRMD_ERR_ARG (-1)      /* invalid argument or shape                          */
#define MRMD_ERR_HIP (-2)      /* HIP runtime error                                  */
#define MRMD_ERR_STATE (-3)    /* call order violated (e.g. update without samples)  */
#define MRMD_ERR_NOMEM (-4)    /* device allocation failed                           */
#define MRMD_ERR_UNSUP (-5)    /* configuration the kernels do not support           */

/* ---- opaque handles ------------------------------------------------------------------ */
typedef struct mrmd_ctx mrmd_ctx;
typedef struct mrmd_atoms mrmd_atoms;
typedef struct mrmd_molecules mrmd_molecules;
typedef struct mrmd_ghost_layer mrmd_ghost_layer;
typedef struct mrmd_cell_list mrmd_cell_list;
typedef struct mrmd_verlet_list mrmd_verlet_list;
typedef struct mrmd_lj mrmd_lj;
typedef struct mrmd_lj_idealgas mrmd_lj_idealgas;
typedef struct mrmd_thermo_force mrmd_thermo_force;
typedef struct mrmd_batch mrmd_batch;

/* ---- plain value types ----------------------------------------------------------------- */
/* periodic box: data::Subdomain(minCorner, maxCorner, ghostLayerThickness)
 * (mrmd/data/Subdomain.hpp:41-62); derived corners are computed exactly as there. */
typedef struct mrmd_box {
    double min[3];
    double max[3];
    double ghost[3];
} mrmd_box;

/* One-position predicate.  Lambdas cannot cross a C ABI; these descriptors cover the
 * predicates the hot path uses (OnePositionPredicate / OneCoordinatePredicate,
 * mrmd/datatypes.hpp:196-210).  c = coordinate `axis` of the position.
 *   MRMD_PRED_ALWAYS          true
 *   MRMD_PRED_NEVER           false
 *   MRMD_PRED_OPEN_RANGE      a < c < b
 *   MRMD_PRED_SYMMETRIC_SLAB  b - tol <= |c - a| <= c_ + tol  (util/IsInSymmetricSlab.hpp:34-50;
 *                             center a, slabMin b, slabMax c, tolerance tol)
 *   MRMD_PRED_SLAB_AT/HY/CG   Slab::isIn{AT,HY,CG}Region with center a, atomistic half
 *                             diameter b, hybrid diameter c (weighting_function/Slab.hpp:44-71)
 *   MRMD_PRED_GREATER         c > a
 *   MRMD_PRED_LESS            c < a
 */
enum {
    MRMD_PRED_ALWAYS = 0,
    MRMD_PRED_NEVER = 1,
    MRMD_PRED_OPEN_RANGE = 2,
    MRMD_PRED_SYMMETRIC_SLAB = 3,
    MRMD_PRED_SLAB_AT = 4,
    MRMD_PRED_SLAB_HY = 5,
    MRMD_PRED_SLAB_CG = 6,
    MRMD_PRED_GREATER = 7,
    MRMD_PRED_LESS = 8
};
typedef struct mrmd_pred1 {
    int32_t kind;
    int32_t axis;
    double a;
    double b;
    double c;
    double tol;
} mrmd_pred1;

/* Two-position predicate op(first(pos_i), second(pos_j)) (TwoPositionsPredicate,
 * mrmd/datatypes.hpp:212-222; the forms used in
 * examples/04_LennardJones_IdealGas_LocalCap/04_LennardJones_IdealGas_LocalCap.cpp:206-227). */
enum { MRMD_PAIR_ALWAYS = 0, MRMD_PAIR_AND = 1, MRMD_PAIR_OR = 2 };
typedef struct mrmd_pred2 {
    int32_t op;
    int32_t pad;
    mrmd_pred1 first;
    mrmd_pred1 second;
} mrmd_pred2;

/* weighting_function::Slab(center, atomisticRegionDiameter, hybridRegionDiameter, nu,
 * interfaceType) (mrmd/weighting_function/Slab.hpp:189-200); interface 0 SMOOTH, 1 ABRUPT */
typedef struct mrmd_slab {
    double center[3];
    double atomistic_diameter;
    double hybrid_diameter;
    int64_t nu;
    int32_t interface_type;
    int32_t pad;
} mrmd_slab;

/* weighting_function::Spherical(center, atomisticRadius, hybridRegionDiameter, exponent)
 * (mrmd/weighting_function/Spherical.hpp:25-99) */
typedef struct mrmd_spherical {
    double center[3];
    double atomistic_radius;
    double hybrid_diameter;
    int64_t exponent;
} mrmd_spherical;

/* ---- field identifiers ------------------------------------------------------------------ */
/* data::Atoms members (mrmd/data/Atoms.hpp:37-46).  Host-side layout of transfers:
 * POS/VEL/FORCE n x 3 doubles (row-major), TYPE int64, MASS/CHARGE/RELMASS doubles. */
enum {
    MRMD_POS = 0,
    MRMD_VEL = 1,
    MRMD_FORCE = 2,
    MRMD_TYPE = 3,
    MRMD_MASS = 4,
    MRMD_CHARGE = 5,
    MRMD_RELMASS = 6
};
/* data::Molecules members (mrmd/data/Molecules.hpp:31-40): POS/FORCE/GRADLAMBDA n x 3
 * doubles, LAMBDA/MODLAMBDA doubles, OFFSET/NUMATOMS int64. */
enum {
    MRMD_MOL_POS = 0,
    MRMD_MOL_FORCE = 1,
    MRMD_MOL_LAMBDA = 2,
    MRMD_MOL_MODLAMBDA = 3,
    MRMD_MOL_GRADLAMBDA = 4,
    MRMD_MOL_OFFSET = 5,
    MRMD_MOL_NUMATOMS = 6
};

/* =========================================================================================
 * context  (replaces Kokkos::ScopeGuard / initialization.cpp:23-44 for this path)
 * ========================================================================================= */
int mrmd_ctx_create(int device, mrmd_ctx** out);
int mrmd_ctx_destroy(mrmd_ctx* ctx);
const char* mrmd_ctx_last_error(const mrmd_ctx* ctx);
/* run on an external stream (e.g. torch.cuda.current_stream().cuda_stream); NULL = own */
int mrmd_ctx_set_stream(mrmd_ctx* ctx, void* hip_stream);
int mrmd_ctx_synchronize(mrmd_ctx* ctx);
/* per-kernel HIP-event timing on the context stream (off by default) */
int mrmd_ctx_set_profiling(mrmd_ctx* ctx, int enable);
/* time only the entry points of this timer name (e.g. "lj_force"); NULL or "" = all */
int mrmd_ctx_set_profile_filter(mrmd_ctx* ctx, const char* kernel);
int mrmd_ctx_kernel_stats(mrmd_ctx* ctx, const char* kernel, int64_t* launches, double* total_ms);
int mrmd_ctx_reset_stats(mrmd_ctx* ctx);
/* bytes of device memory currently held by this context's objects */
int mrmd_ctx_device_bytes(mrmd_ctx* ctx, int64_t* bytes);

/* =========================================================================================
 * containers
 * ========================================================================================= */
/* data::Atoms(numAtoms) (Atoms.hpp:123-133): all fields zero, numLocal = numGhost = 0 */
int mrmd_atoms_create(mrmd_ctx* ctx, int64_t size, mrmd_atoms** out);
int mrmd_atoms_destroy(mrmd_atoms* atoms);
/* Atoms::resize (Atoms.hpp:89-93): keeps the first min(old,new) records */
int mrmd_atoms_resize(mrmd_atoms* atoms, int64_t size);
int mrmd_atoms_get_counts(const mrmd_atoms* atoms, int64_t* nlocal, int64_t* nghost, int64_t* size);
/* numLocalAtoms / numGhostAtoms (Atoms.hpp:120-121) */
int mrmd_atoms_set_counts(mrmd_atoms* atoms, int64_t nlocal, int64_t nghost);
/* data::deep_copy host <-> device (Atoms.hpp:150-159) for one field, records [offset, offset+n) */
int mrmd_atoms_upload(mrmd_atoms* atoms, int field, const void* host, int64_t offset, int64_t n);
int mrmd_atoms_download(const mrmd_atoms* atoms, int field, void* host, int64_t offset, int64_t n);
/* Atoms::setForce / Cabana::deep_copy(slice, value) over all records (Atoms.hpp:72) */
int mrmd_atoms_fill(mrmd_atoms* atoms, int field, double value);
/* device pointer of a field; vector fields are component-major with the returned stride
 * (x at ptr[i], y at ptr[stride+i], z at ptr[2*stride+i]); POS is 4 doubles per record
 * (x, y, z, type) and returns stride 0.  Pointers change when the container grows. */
int mrmd_atoms_device_ptr(mrmd_atoms* atoms, int field, void** ptr, int64_t* stride);

int mrmd_molecules_create(mrmd_ctx* ctx, int64_t size, mrmd_molecules** out);
int mrmd_molecules_destroy(mrmd_molecules* mols);
int mrmd_molecules_resize(mrmd_molecules* mols, int64_t size);
int mrmd_molecules_get_counts(const mrmd_molecules* mols, int64_t* nlocal, int64_t* nghost, int64_t* size);
int mrmd_molecules_set_counts(mrmd_molecules* mols, int64_t nlocal, int64_t nghost);
int mrmd_molecules_upload(mrmd_molecules* mols, int field, const void* host, int64_t offset, int64_t n);
int mrmd_molecules_download(const mrmd_molecules* mols, int field, void* host, int64_t offset, int64_t n);
/* Molecules::setForce (Molecules.hpp:60) */
int mrmd_molecules_fill(mrmd_molecules* mols, int field, double value);
/* data::createMoleculeForEachAtom (mrmd/data/MoleculesFromAtoms.cpp:19-39) */
int mrmd_molecules_for_each_atom(mrmd_molecules* mols, const mrmd_atoms* atoms);

/* =========================================================================================
 * periodic halo  (communication::GhostLayer, mrmd/communication/GhostLayer.hpp:28-56)
 * ========================================================================================= */
int mrmd_ghost_layer_create(mrmd_ctx* ctx, mrmd_ghost_layer** out);
int mrmd_ghost_layer_destroy(mrmd_ghost_layer* gl);
/* GhostLayer::exchangeRealAtoms -> PeriodicMapping::mapIntoDomain (PeriodicMapping.cpp:30-58) */
int mrmd_halo_wrap(mrmd_atoms* atoms, const mrmd_box* box);
/* GhostLayer::createGhostAtoms -> GhostExchange::createGhostAtomsXYZ (GhostExchange.cpp:171-181):
 * ordered ghost creation, X then Y then Z; returns the new ghost count */
int mrmd_halo_create(mrmd_ghost_layer* gl, mrmd_atoms* atoms, const mrmd_box* box, int64_t* nghost);
/* GhostExchange::createGhostAtoms for one axis (GhostExchange.cpp:59-169); reset != 0 first
 * runs resetCorrespondingRealAtoms (GhostExchange.cpp:183-187) */
int mrmd_halo_create_axis(mrmd_ghost_layer* gl, mrmd_atoms* atoms, const mrmd_box* box, int axis, int reset);
/* GhostLayer::updateGhostAtoms -> UpdateGhostAtoms::updateOnlyPos (UpdateGhostAtoms.cpp:31-68) */
int mrmd_halo_update(mrmd_ghost_layer* gl, mrmd_atoms* atoms, const mrmd_box* box);
/* GhostLayer::contributeBackGhostToReal -> AccumulateForce::ghostToReal (AccumulateForce.cpp:25-47) */
int mrmd_halo_fold_forces(mrmd_ghost_layer* gl, mrmd_atoms* atoms);
/* correspondingRealAtom view (GhostExchange.hpp:39-40) */
int mrmd_halo_get_corresponding(const mrmd_ghost_layer* gl, int64_t* host, int64_t n);
int mrmd_halo_set_corresponding(mrmd_ghost_layer* gl, const int64_t* host, int64_t n);
/* MultiResGhostLayer::exchangeRealAtoms -> realAtomsExchange (MultiResRealAtomsExchange.cpp:23-73) */
int mrmd_multires_wrap(mrmd_molecules* mols, mrmd_atoms* atoms, const mrmd_box* box);
/* MultiResGhostLayer::createGhostAtoms -> MultiResPeriodicGhostExchange::createGhostAtomsXYZ
 * (MultiResPeriodicGhostExchange.cpp:250-265) */
int mrmd_multires_halo_create(mrmd_ghost_layer* gl, mrmd_molecules* mols, mrmd_atoms* atoms, const mrmd_box* box);
/* MultiResPeriodicGhostExchange::createGhostAtoms, one axis (MultiResPeriodicGhostExchange.cpp:60-248) */
int mrmd_multires_halo_create_axis(mrmd_ghost_layer* gl, mrmd_molecules* mols, mrmd_atoms* atoms,
                                   const mrmd_box* box, int axis, int reset);

/* =========================================================================================
 * neighbour search  (Cabana::LinkedCellList and Cabana::VerletList<HalfNeighborTag,
 * VerletLayout2D, TeamOpTag>, typedefs mrmd/datatypes.hpp:179-194)
 * ========================================================================================= */
int mrmd_cell_list_create(mrmd_ctx* ctx, mrmd_cell_list** out);
int mrmd_cell_list_destroy(mrmd_cell_list* cl);
/* LinkedCellList(pos, begin, end, delta[3], min[3], max[3]) (tests/NVT/NVT.cpp:138-143):
 * in-cell order is ascending particle index */
int mrmd_cell_list_build(mrmd_cell_list* cl, const mrmd_atoms* atoms, int64_t begin, int64_t end,
                         const double* delta, const double* grid_min, const double* grid_max);
int mrmd_cell_list_dims(const mrmd_cell_list* cl, int32_t* dims);
/* binning data: cell of particle begin+k, counts and offsets per cell, permutation
 * (sorted slot k -> particle index); any pointer may be NULL */
int mrmd_cell_list_download(const mrmd_cell_list* cl, int32_t* cell_of, int32_t* counts,
                            int64_t* offsets, int64_t* permute);
/* Atoms::permute(LinkedCellList&) -> Cabana::permute (Atoms.hpp:95) */
int mrmd_atoms_permute(mrmd_atoms* atoms, const mrmd_cell_list* cl);

int mrmd_verlet_create(mrmd_ctx* ctx, mrmd_verlet_list** out);
int mrmd_verlet_destroy(mrmd_verlet_list* vl);
/* HalfVerletList::build(pos, begin, end, neighborhoodRadius, cellRatio, gridMin, gridMax,
 * maxNeigh) (examples/02_LennardJones_NVE/02_LennardJones_NVE.cpp:156-163): binning of every
 * record (local and ghost), 27-cell stencil with Cabana's min-distance prune, d^2 <= r^2.
 * The half list (Cabana's position-lexicographic discriminator) is exact; the device keeps
 * both directions so force kernels can run without atomics, in tiles of cells whose
 * neighbourhood positions the force kernel stages in LDS. */
int mrmd_verlet_build(mrmd_verlet_list* vl, const mrmd_atoms* atoms, int64_t begin, int64_t end,
                      double r_n, double cell_ratio, const double* grid_min, const double* grid_max,
                      int64_t max_neigh);
/* the same over molecule centres (LJ_IdealGas molecule list, LJ_IdealGas.test.cpp:108-120) */
int mrmd_verlet_build_molecules(mrmd_verlet_list* vl, const mrmd_molecules* mols, int64_t begin,
                                int64_t end, double r_n, double cell_ratio, const double* grid_min,
                                const double* grid_max, int64_t max_neigh);
/* HalfNeighborList::numNeighbor for every record (counts of the half list, zero outside
 * [begin, end)); total = sum of counts (tests/LennardJones/LennardJones.cpp:72-80) */
int mrmd_verlet_half_counts(mrmd_verlet_list* vl, int32_t* counts, int64_t n, int64_t* total);
/* HalfNeighborList::getNeighbor rows [0, n) of the half list into host[row * row_stride + k] */
int mrmd_verlet_half_neighbors(mrmd_verlet_list* vl, int32_t* host, int64_t n, int64_t row_stride);
/* both-direction rows (record begin+r in row r), same layout; *max_count is the widest row */
int mrmd_verlet_full_neighbors(mrmd_verlet_list* vl, int32_t* host, int32_t* counts, int64_t n,
                               int64_t row_stride);
int mrmd_verlet_info(const mrmd_verlet_list* vl, int64_t* rows, int32_t* max_full, int64_t* total_full,
                     int32_t* dims);
/* pairs of the list closer than r at the atoms' current positions: half-list pairs and
 * both-direction entries (the algorithmic pair counts behind the force-kernel FLOP figure;
 * no reference counterpart) */
int mrmd_verlet_count_within(mrmd_verlet_list* vl, const mrmd_atoms* atoms, double r, int64_t* half_pairs,
                             int64_t* full_entries);
/* Launch shape of the tiled list and the tiled force kernel (no reference counterpart; the
 * reference's Cabana list has no tiling): a tile is tile_cells cells along z (0 = chosen
 * from the density at the next build) in tile_columns x tile_columns cell columns (1-3;
 * 1 when the cell ratio is below 1); block = workgroup size (64, 128, 192 or 256).
 * Applies from the next build; results do not depend on it. */
int mrmd_verlet_set_tiling(mrmd_verlet_list* vl, int tile_cells, int tile_columns, int block);
/* No counterpart in the reference (a test knob): on != 0 builds the list with the 64-bit
 * entry-offset form of the build kernel, which otherwise serves only lists of 2^31 entries
 * or more (about 20M atoms); the list is the same either way. */
int mrmd_verlet_set_wide_offsets(mrmd_verlet_list* vl, int on);
/* No counterpart in the reference: the most records a tile may stage in LDS (1..2048; 0 =
 * 2048).  The build shrinks the tile height, then its columns, until every tile fits; when
 * even one cell column by one cell does not (dense inputs), each tile stages its
 * neighbourhood in a slice of global memory that the list build and force kernels use in
 * place of LDS (same lists, same bits).  A small limit forces that path in tests. */
int mrmd_verlet_set_stage_limit(mrmd_verlet_list* vl, int limit);
/* the tiling in use: cells per tile, columns, workgroup size, largest staged neighbourhood */
int mrmd_verlet_get_tiling(const mrmd_verlet_list* vl, int* tile_cells, int* tile_columns, int* block,
                           int* max_staged);

/* =========================================================================================
 * Lennard-Jones  (action::LennardJones, mrmd/action/LennardJones.hpp:98-206)
 * ========================================================================================= */
/* CappedLennardJonesPotential(cap[], rc[], sigma[], eps[], numTypes, isShifted)
 * (LennardJones.cpp:81-116, 61-79).  type_stride selects the pair index
 * type_i * type_stride + type_j: LennardJones hard-codes 1 (LennardJones.cpp:52),
 * LJ_IdealGas uses numTypes (LJ_IdealGas.cpp:139). */
int mrmd_lj_create(mrmd_ctx* ctx, const double* cap, const double* rc, const double* sigma,
                   const double* epsilon, int64_t num_types, int shifted, int64_t type_stride, mrmd_lj** out);
int mrmd_lj_destroy(mrmd_lj* lj);
/* precomputed table, num_types^2 x 10 doubles in PrecomputedValues order (LennardJones.hpp:27-39) */
int mrmd_lj_table(const mrmd_lj* lj, double* host);
/* computeForceAndEnergy(distSqr, typeIdx) (LennardJones.hpp:52-78) on n host values, on the device */
int mrmd_lj_eval(mrmd_lj* lj, const double* dist_sqr, const int64_t* type_idx, int64_t n,
                 double* force_factor, double* energy);
/* LennardJones::apply / apply_if (LennardJones.hpp:136-206): adds pair forces to local atoms.
 * pred == NULL is apply().  Energy and virial stay on the device until requested. */
int mrmd_lj_apply(mrmd_lj* lj, mrmd_atoms* atoms, mrmd_verlet_list* vl, const mrmd_pred2* pred);
/* No counterpart in the reference: mrmd_lj_apply (no predicate) for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Adds the pair forces to the local atoms with the same kernel and bits
 * as mrmd_lj_apply and writes energy and virial to the caller's device memory
 * energy_virial_dev[0..1] (and to the table, for a later mrmd_lj_energy_virial), without
 * synchronising.  It calls no hipMalloc / hipFree: the per-workgroup partials go to a static
 * device array of MRMD_LJ_NOALLOC_MAX_TILES x 2 doubles (1 MiB) that each GPU the library
 * runs on holds until the process ends and that every call on that GPU shares, so calls that
 * overlap on one GPU race.  Symmetric tables only (the full-row kernel); a list with more
 * tiles than MRMD_LJ_NOALLOC_MAX_TILES is refused with MRMD_ERR_UNSUP. */
#define MRMD_LJ_NOALLOC_MAX_TILES 65536
int mrmd_lj_apply_noalloc(mrmd_lj* lj, mrmd_atoms* atoms, mrmd_verlet_list* vl, double* energy_virial_dev);
/* getEnergy / getVirial (LennardJones.cpp:35-36); synchronises */
int mrmd_lj_energy_virial(mrmd_lj* lj, double* energy, double* virial);
/* Adjoint of the force map F(x; eps, sigma) (no reference counterpart; for force matching):
 * given upstream g = dL/dF for local atoms (n x 3), returns dL/dx (n x 3, local atoms) and
 * dL/d(eps_ab), dL/d(sigma_ab) per type pair (num_types^2 each, symmetric: entry (a,b) =
 * entry (b,a) = derivative w.r.t. the shared pair parameter).  Pair terms are recomputed,
 * never stored.  gl (the ghost layer that created the ghosts) is required when ghosts
 * exist: a ghost partner takes the upstream gradient of its real atom.  device != 0: dLdF
 * and dLdx are device pointers (the parameter gradients are always host arrays). */
int mrmd_lj_adjoint(mrmd_lj* lj, const mrmd_atoms* atoms, mrmd_verlet_list* vl, const mrmd_ghost_layer* gl,
                    const double* dLdF, double* dLdx, double* dLdeps, double* dLdsigma, int device);
/* reparameterise an existing table in place (same cap/rc/shift flags) */
int mrmd_lj_set_params(mrmd_lj* lj, const double* sigma, const double* epsilon);

/* =========================================================================================
 * integrators  (action::VelocityVerlet, VelocityVerletLangevinThermostat)
 * ========================================================================================= */
/* VelocityVerlet::preForceIntegrate (VelocityVerlet.cpp:26-67); returns sqrt(max |dx|^2) in
 * *max_disp (synchronises); max_disp may be NULL (no synchronisation) */
int mrmd_vv_pre(mrmd_atoms* atoms, double dt, double* max_disp);
/* VelocityVerlet::postForceIntegrate (VelocityVerlet.cpp:69-90) */
int mrmd_vv_post(mrmd_atoms* atoms, double dt);
/* VelocityVerletLangevinThermostat::preForceIntegrate_apply_if
 * (VelocityVerletLangevinThermostat.hpp:63-136).  Noise: Philox4x32-10 keyed by seed,
 * counter (call, atom index) - replaces Random_XorShift1024_Pool(1234). pred NULL = always. */
int mrmd_langevin_pre(mrmd_atoms* atoms, double dt, double zeta, double temperature, uint64_t seed,
                      uint64_t call, const mrmd_pred1* pred, double* max_disp);
/* BerendsenThermostat::apply (mrmd/action/BerendsenThermostat.cpp:27-50): local velocities
 * scaled by sqrt(1 + gamma (target / current - 1)); nothing happens for current <= 0, and a
 * target <= 0 is an error */
int mrmd_berendsen_apply(mrmd_atoms* atoms, double current_temperature, double target_temperature, double gamma);

/* The time loop of the reference's LJ examples over the calls above, run natively:
 * examples/02_LennardJones_NVE/02_LennardJones_NVE.cpp:135-186 (velocity Verlet; exchange,
 * ghosts and list on a rebuild, ghost update otherwise) with the Langevin pre-force
 * integrator of examples/01_LennardJones_NVT.cpp:139-214 (langevin = 1) and, when `cells` is
 * not NULL, LinkedCellList + Atoms::permute on every rebuild (tests/NVT/NVT.cpp:138-144).
 * grid_min / grid_max (with cell_ratio) give the Verlet list's binning grid, the box's
 * corners the cell list (delta r_n).
 * The state fields are read and updated, so consecutive calls continue one run; the
 * results equal those of the same calls made one by one, bit for bit (a step's post-force
 * and the next step's pre-force integration run as one pass with the same arithmetic). */
typedef struct mrmd_md_loop {
    double dt;
    double neighbor_radius; /* r_n = rc + skin of the list builds */
    double skin;            /* rebuild when the accumulated displacement reaches skin / 2 */
    double cell_ratio;
    double grid_min[3];     /* the list's binning grid (HalfVerletList::build gridMin/gridMax) */
    double grid_max[3];
    int64_t max_neigh;
    int32_t langevin;       /* 0: VelocityVerlet, 1: VelocityVerletLangevinThermostat */
    int32_t reserved;
    double zeta, temperature;
    uint64_t seed;
    uint64_t call;          /* state: Langevin noise counter, one per step */
    double max_disp;        /* state: displacement accumulated since the last rebuild */
    int64_t rebuilds;       /* state */
    int64_t steps;          /* state */
} mrmd_md_loop;
int mrmd_md_run(mrmd_atoms* atoms, mrmd_ghost_layer* gl, mrmd_verlet_list* vl, mrmd_lj* lj,
                mrmd_cell_list* cells, const mrmd_box* box, mrmd_md_loop* loop, int64_t nsteps);
/* analysis::getKineticEnergy (mrmd/analysis/KineticEnergy.hpp:26-39); synchronises */
int mrmd_kinetic_energy(const mrmd_atoms* atoms, double* energy);

/* =========================================================================================
 * AdResS  (weighting_function::Slab, UpdateMolecules, LJ_IdealGas,
 *          ContributeMoleculeForceToAtoms, ThermodynamicForce)
 * ========================================================================================= */
/* Slab::operator() (Slab.hpp:161-187) on n host x values: out5[i] = lambda, modLambda, grad */
int mrmd_slab_eval(mrmd_ctx* ctx, const mrmd_slab* slab, const double* x, int64_t n, double* out5);
/* Spherical::operator() (Spherical.hpp:37-76) on n host points (n x 3, row-major):
 * out4[4 i ..] = lambda, d lambda / dx, dy, dz */
int mrmd_spherical_eval(mrmd_ctx* ctx, const mrmd_spherical* s, const double* xyz, int64_t n, double* out4);
/* UpdateMolecules::update (mrmd/action/UpdateMolecules.hpp:24-70); slab NULL = COM only */
int mrmd_molecules_update(mrmd_molecules* mols, const mrmd_atoms* atoms, const mrmd_slab* slab);
/* ContributeMoleculeForceToAtoms::update (ContributeMoleculeForceToAtoms.cpp:23-48) */
int mrmd_molecules_contribute_force(const mrmd_molecules* mols, mrmd_atoms* atoms);

/* LJ_IdealGas(cap[], rc[], sigma[], eps[], numTypes, doShift) (LJ_IdealGas.cpp:271-292) */
int mrmd_ljig_create(mrmd_ctx* ctx, const double* cap, const double* rc, const double* sigma,
                     const double* epsilon, int64_t num_types, int shifted, mrmd_lj_idealgas** out);
int mrmd_ljig_destroy(mrmd_lj_idealgas* s);
/* setCompensationEnergySamplingInterval / setCompensationEnergyUpdateInterval (LJ_IdealGas.hpp:73-80) */
int mrmd_ljig_set_intervals(mrmd_lj_idealgas* s, int64_t sampling, int64_t update);
/* Tuning of the tiled form (no counterpart in the reference): pairs whose partner is not a
 * row (a ghost image) deposit the partner side from a per-workgroup queue of at most
 * `capacity` pairs, 0..4096 (default 4096); rows whose pairs do not fit deposit their own.
 * Results do not depend on it beyond the order of FP64 atomic sums. */
int mrmd_ljig_set_queue(mrmd_lj_idealgas* s, int capacity);
/* LJ_IdealGas::run (LJ_IdealGas.cpp:227-260); synchronises for the energy.  Lists built
 * over molecules of one atom each with a single-type table run the tiled full-row form;
 * otherwise the half-list form. */
int mrmd_ljig_run(mrmd_lj_idealgas* s, mrmd_molecules* mols, mrmd_verlet_list* vl, mrmd_atoms* atoms,
                  double* energy);
/* which: 0 compensation energy, 1 counter, 2 mean (getMeanCompensationEnergy), 200 x numTypes */
int mrmd_ljig_get_histogram(const mrmd_lj_idealgas* s, int which, double* host);
int mrmd_ljig_set_histogram(mrmd_lj_idealgas* s, int which, const double* host);
/* number of run() calls so far, which decides the sampling and update steps of the
 * compensation energy (runCounter_, LJ_IdealGas.cpp:241-257): get (if not NULL) reads it
 * before set (if not NULL) replaces it; for checkpoints */
int mrmd_ljig_run_counter(mrmd_lj_idealgas* s, int64_t* get, const int64_t* set);

/* ThermodynamicForce(targetDensity[], subdomain, binWidth, modulation[], enforceSymmetry,
 * usePeriodicity) (ThermodynamicForce.cpp:25-57) */
int mrmd_tf_create(mrmd_ctx* ctx, const double* target_density, const double* modulation,
                   int64_t num_types, const mrmd_box* box, double bin_width, int enforce_symmetry,
                   int use_periodicity, mrmd_thermo_force** out);
int mrmd_tf_destroy(mrmd_thermo_force* tf);
int mrmd_tf_num_bins(const mrmd_thermo_force* tf, int64_t* nbins);
/* ThermodynamicForce::sample (ThermodynamicForce.cpp:74-86) */
int mrmd_tf_sample(mrmd_thermo_force* tf, const mrmd_atoms* atoms);
/* update / update_if (ThermodynamicForce.hpp:124-152); pred over the bin centre, NULL = always */
int mrmd_tf_update(mrmd_thermo_force* tf, double smoothing_sigma, double smoothing_range,
                   const mrmd_pred1* pred);
/* apply / apply_if (ThermodynamicForce.hpp:98-122) */
int mrmd_tf_apply(const mrmd_thermo_force* tf, mrmd_atoms* atoms, const mrmd_pred1* pred);
/* applyInterpolated_if (ThermodynamicForce.hpp:154-216) */
int mrmd_tf_apply_interpolated(const mrmd_thermo_force* tf, mrmd_atoms* atoms, const mrmd_pred1* pred);
/* which: 0 force (getForce), 1 density profile (getDensityProfile); nbins x numTypes */
int mrmd_tf_get(const mrmd_thermo_force* tf, int which, double* host);
/* setForce (ThermodynamicForce.hpp:55) */
int mrmd_tf_set_force(mrmd_thermo_force* tf, const double* host);
/* getNumberOfDensityProfileSamples (ThermodynamicForce.hpp:63) */
int mrmd_tf_samples(const mrmd_thermo_force* tf, int64_t* samples);
/* restore the accumulated density profile (nbins x numTypes) and its sample count, the
 * state sample() builds between updates (ThermodynamicForce.cpp:74-86); for checkpoints */
int mrmd_tf_set_density(mrmd_thermo_force* tf, const double* host, int64_t samples);
/* getMuLeft / getMuRight (ThermodynamicForce.cpp:98-130) */
int mrmd_tf_mu(const mrmd_thermo_force* tf, double* left, double* right);

/* The LJ / ideal-gas AdResS slab step over the calls above, run natively: Langevin
 * pre-force integration; on a rebuild molecule centres, MultiResPeriodicGhostExchange
 * (exchange + ghosts) and the molecule Verlet list (grid box - r_n .. box + r_n, cell ratio
 * 1), otherwise a ghost refresh; UpdateMolecules with the slab weighting; zero atom and
 * molecule forces; LJ_IdealGas::run (LJ_IdealGas.cpp:227-260);
 * ContributeMoleculeForceToAtoms; ghost fold; ThermodynamicForce::sample, update_if every
 * tf_every steps (ThermodynamicForce.hpp:124-152), apply; post-force integration.  State
 * fields are read and updated, so consecutive calls continue one run; results equal those
 * of the same calls made one by one. */
typedef struct mrmd_adress_loop {
    double dt;
    double neighbor_radius;
    double skin;
    double zeta, temperature;
    uint64_t seed;
    int64_t tf_every;      /* thermodynamic-force update interval in steps, 0 = never */
    double tf_sigma, tf_range;
    mrmd_pred1 tf_pred;    /* update_if predicate over the bin centres */
    uint64_t call;         /* state: Langevin noise counter */
    double max_disp;       /* state */
    int64_t rebuilds;      /* state */
    int64_t steps;         /* state: steps taken (the update interval counts them) */
    double energy;         /* out: LJ_IdealGas energy of the last step */
} mrmd_adress_loop;
int mrmd_adress_md_run(mrmd_atoms* atoms, mrmd_molecules* mols, mrmd_ghost_layer* gl, mrmd_verlet_list* vl,
                       mrmd_lj_idealgas* ig, mrmd_thermo_force* tf, const mrmd_slab* slab, const mrmd_box* box,
                       mrmd_adress_loop* loop, int64_t nsteps);

/* MultiHistogram operations on host arrays [nbins x nhist], executed on the device
 * (MultiHistogram.cpp:113-212, MultiHistogram.cpp:76-90) */
int mrmd_hist_gradient(mrmd_ctx* ctx, double min, double max, int64_t nbins, int64_t nhist,
                       const double* in, int periodic, double* out);
int mrmd_hist_smoothen(mrmd_ctx* ctx, double min, double max, int64_t nbins, int64_t nhist,
                       const double* in, double sigma, double range, int periodic, double* out);

/* =========================================================================================
 * ensemble engine (no reference counterpart; for thermodynamic-force iterations and
 * parameter sweeps): many independent periodic replicas of uneven size advanced together.
 * Every phase of a step is one launch over all replicas; blocks of steps are captured once
 * as a hipGraph and replayed.  A replica is either plain LJ under the Langevin thermostat
 * (the loop of examples/01_LennardJones_NVT/01_LennardJones_NVT.cpp:139-214) or a
 * monatomic LJ / ideal-gas AdResS slab along x (LJ_IdealGas.cpp:52-260 with one molecule
 * per atom, Slab.hpp:96-187, ThermodynamicForce.hpp:98-152).  Each replica's result does
 * not depend on the other replicas: a replica run alone gives the same bits.
 * ========================================================================================= */
typedef struct mrmd_replica {
    int64_t num_atoms;
    double box[3];               /* edges; the periodic box spans [0, box) */
    double dt;
    double gamma;                /* Langevin friction */
    double temperature;
    double mass;
    uint64_t seed;               /* Langevin noise key */
    int32_t adress;              /* 0 plain LJ, 1 AdResS slab centred in the box */
    int32_t tf_every;            /* thermodynamic-force update interval in steps (0: never) */
    double atomistic_diameter;   /* slab */
    double hybrid_diameter;
    int64_t nu;
    double tf_bin_width;         /* thermodynamic force on x-bins */
    double tf_target_density;
    double tf_sigma;             /* smoothing width and range (MultiHistogram::smoothen) */
    double tf_range;
    double tf_modulation;
    int32_t tf_symmetric;
    int32_t tf_periodic;
} mrmd_replica;

int mrmd_batch_create(mrmd_ctx* ctx, const mrmd_replica* replicas, int64_t num_replicas, double rc, double skin,
                      mrmd_batch** out);
int mrmd_batch_destroy(mrmd_batch* b);
/* one LJ table for all replicas: capping distance, sigma, epsilon, energy shift at rc */
int mrmd_batch_set_lj(mrmd_batch* b, double cap, double sigma, double epsilon, int shifted);
/* launch shape: target number of records a workgroup stages (sets each replica's tile
 * height), workgroup size (64, 128, 192 or 256); forces a list rebuild at the next step */
int mrmd_batch_set_tiling(mrmd_batch* b, int target_staged, int block);
int mrmd_batch_get_tiling(const mrmd_batch* b, int* target_staged, int* block, int32_t* tile_cells);
/* upper bound on a workgroup's LDS stage area in records (0 = sized from the expected
 * neighbourhood, the default); tiles whose neighbourhood is larger read their records from
 * global memory with identical arithmetic, so results do not depend on it (a tuning and
 * testing knob, no reference counterpart); forces a list rebuild at the next step */
int mrmd_batch_set_stage_cap(mrmd_batch* b, int limit);
/* all replicas' atoms concatenated in replica order, each replica's atoms in the order
 * given, n_total x 3 row-major (POS, VEL; FORCE for download) */
int mrmd_batch_upload(mrmd_batch* b, int field, const double* host);
int mrmd_batch_download(mrmd_batch* b, int field, double* host);
/* advance every replica by nsteps; graph_steps > 0 replays a captured graph of that many
 * steps (a remainder runs as plain launches), 0 launches every kernel directly */
int mrmd_batch_run(mrmd_batch* b, int64_t nsteps, int graph_steps);
/* per-replica potential energy of the last force evaluation and current kinetic energy */
int mrmd_batch_energies(mrmd_batch* b, double* potential, double* kinetic);
/* per-replica list rebuild counts; steps taken */
int mrmd_batch_stats(mrmd_batch* b, int64_t* rebuilds, int64_t* steps);
/* thermodynamic force (which 0) or density profile (which 1) of one replica's x-bins */
int mrmd_batch_tf_get(mrmd_batch* b, int64_t replica, int which, double* host, int64_t n);
/* Checkpoint of the ensemble's dynamic state (no counterpart in the reference, whose
 * checkpoints are H5MD files of one system, io/DumpH5MD.hpp:30-75, io/RestoreH5MD.hpp:36-39):
 * positions, velocities, forces and each atom's given index (atoms sit in cell order on the
 * device), the step counter that keys the Langevin noise, rebuild
 * and density-sample counts, thermodynamic-force and density bins, compensation-energy
 * sums, counts and means.  state_bytes gives the blob's size; save writes it, load reads a
 * blob of the same replicas into an ensemble created with the same replica descriptions.
 * Both end the current list epoch (every replica rebuilds at its next step), so the run
 * that continues after a save and the run restored from it stay bit-identical. */
int mrmd_batch_state_bytes(const mrmd_batch* b, int64_t* bytes);
int mrmd_batch_state_save(mrmd_batch* b, void* host, int64_t bytes);
int mrmd_batch_state_load(mrmd_batch* b, const void* host, int64_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* MRMD_HIP_H */
