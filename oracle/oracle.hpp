// TEST INFRASTRUCTURE ONLY - NOT PART OF THE PRODUCT.
//
// CPU restatement of MRMD's per-timestep hot path (XzzX/mrmd v0.2.0), written from the
// reference's semantics and used exclusively as the parity checker by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  Nothing in mrmd_amd/ may
// include, link or call anything in this directory.
//
// Layout follows the reference on purpose (it is also the timed "host-core reference"):
//  * atoms are an array of 104-byte records {pos[3], vel[3], force[3], type, mass, charge,
//    relMass} = Cabana AoSoA with VECTOR_LENGTH=1 (mrmd/data/Atoms.hpp:37-54);
//  * the half Verlet list is a 2-D (counts[N], neighbors[N][maxN]) layout with Cabana's
//    position-lexicographic half discriminator;
//  * force loops are parallel over i with atomic F_j updates (LennardJones.hpp:158-205).
//
// All arithmetic is compiled with -ffp-contract=off so that every expression is evaluated
// exactly as written (no fused multiply-add), which is the convention the HIP library
// follows for every comparison that must be bit-exact.
#pragma once

#include <cstdint>
#include <vector>

extern "C" {
/// periodic box: min/max corner and ghost-layer thickness per dimension
typedef struct
{
    double min[3];
    double max[3];
    double ghost[3];
} or_box;

/// one-position predicate descriptor (kinds listed in oracle.cpp: evalPred1)
typedef struct
{
    int32_t kind;
    int32_t axis;
    double a;
    double b;
    double c;
    double tol;
} or_pred1;

/// two-position predicate: op(first(pos_i), second(pos_j))
typedef struct
{
    int32_t op;
    int32_t pad;
    or_pred1 first;
    or_pred1 second;
} or_pred2;

/// weighting_function::Slab parameters (mrmd/weighting_function/Slab.hpp:189-200)
typedef struct
{
    double center[3];
    double atomistic_diameter;
    double hybrid_diameter;
    int64_t nu;
    int32_t interface_type;  ///< 0 SMOOTH, 1 ABRUPT
    int32_t pad;
} or_slab;

/// parameters of the built-in time loops (examples/01 and 02 of the reference)
typedef struct
{
    double dt;
    double skin;
    double rc;
    double sigma;
    double epsilon;
    double cap;
    double cell_ratio;
    double gamma;
    double temperature;
    uint64_t seed;
    int64_t max_neigh;
    int64_t nsteps;
    int64_t energy_every;
    int32_t shifted;
    int32_t langevin;
    int32_t permute;
    int32_t pad;
} or_md_params;
}
