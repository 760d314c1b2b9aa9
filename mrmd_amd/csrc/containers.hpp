// Device layouts of data::Atoms and data::Molecules (mrmd/data/Atoms.hpp:33-177,
// mrmd/data/Molecules.hpp:27-176), redesigned for HBM streaming and neighbour gathers:
//   pos    : double4 {x, y, z, type} per record - one 32-byte aligned read per neighbour
//            gather; the type travels with the position so force kernels need no
//            second gather
//   vel    : component-major SoA, vel[d * cap + i]
//   force  : component-major SoA, force[d * cap + i]
//   mass, charge, relMass : SoA
// Records [0, nlocal) are local atoms, [nlocal, nlocal + nghost) ghosts (the reference's
// "HaloAtoms" segment).
#pragma once

#include "common.hpp"

namespace mrmd_hip
{
struct AtomsView
{
    double4* pos;
    double* vel;
    double* force;
    double* mass;
    double* charge;
    double* relmass;
    int64_t cap;
};

struct MolView
{
    double4* pos;  ///< {x, y, z, unused}
    double* force;  ///< [3][cap]
    double* lambda;
    double* modLambda;
    double* grad;  ///< [3][cap]
    int64_t* offset;
    int64_t* numAtoms;
    int64_t cap;
};
}  // namespace mrmd_hip

struct mrmd_atoms
{
    mrmd_hip::CtxPtr ctx;
    int64_t nlocal = 0, nghost = 0, size = 0, cap = 0;
    int64_t maxType = 0;  ///< largest type ever uploaded (kernels index tables with it)
    /// every record [0, size) has mass massValue (tracked by the host-side writers; lets the
    /// Langevin step compute its mass-dependent factors once)
    bool massUniform = true;
    double massValue = 0.0;
    /// the same for charge and relmass (a permutation leaves a uniform field as it is)
    bool chargeUniform = true, relmassUniform = true;
    double chargeValue = 0.0, relmassValue = 0.0;
    mrmd_hip::DBuf<double4> pos;
    mrmd_hip::DBuf<double> vel, force, mass, charge, relmass;
    mrmd_hip::DBuf<double> stage;  ///< transfer staging
    /// second set of field buffers: permutation gathers into it and swaps it in
    mrmd_hip::DBuf<double4> sparePos;
    mrmd_hip::DBuf<double> spareVel, spareForce, spareMass, spareCharge, spareRelmass;

    mrmd_hip::AtomsView view() const
    {
        return {pos.p, vel.p, force.p, mass.p, charge.p, relmass.p, cap};
    }
    mrmd_hip::AtomsView spareView() const
    {
        return {sparePos.p, spareVel.p, spareForce.p, spareMass.p, spareCharge.p, spareRelmass.p, cap};
    }
    /// grow the capacity to at least `need` records, preserving [0, size)
    void ensureCap(int64_t need);
    /// set size (records), growing when needed; new records are zero
    void setSize(int64_t n);
};

struct mrmd_molecules
{
    mrmd_hip::CtxPtr ctx;
    int64_t nlocal = 0, nghost = 0, size = 0, cap = 0;
    mrmd_hip::DBuf<double4> pos;
    mrmd_hip::DBuf<double> force, lambda, modLambda, grad;
    mrmd_hip::DBuf<int64_t> offset, numAtoms;
    mrmd_hip::DBuf<double> stage;
    /// bumped by every call that may change the record count or an offset / atom count
    /// (lists built over the molecules remember the value they saw)
    uint64_t layoutVersion = 0;

    mrmd_hip::MolView view() const
    {
        return {pos.p, force.p, lambda.p, modLambda.p, grad.p, offset.p, numAtoms.p, cap};
    }
    void ensureCap(int64_t need);
    void setSize(int64_t n);
};

namespace mrmd_hip
{
// copy record src -> dst (all fields) with a shift of `shift` on axis `axis`
__device__ inline void copyAtom(const AtomsView& a, int64_t dst, int64_t src, int axis, double shift)
{
    double4 p = a.pos[src];
    if (axis == 0) p.x += shift;
    if (axis == 1) p.y += shift;
    if (axis == 2) p.z += shift;
    a.pos[dst] = p;
    for (int d = 0; d < 3; ++d)
    {
        a.vel[d * a.cap + dst] = a.vel[d * a.cap + src];
        a.force[d * a.cap + dst] = a.force[d * a.cap + src];
    }
    a.mass[dst] = a.mass[src];
    a.charge[dst] = a.charge[src];
    a.relmass[dst] = a.relmass[src];
}

// PeriodicMapping::mapIntoDomain (PeriodicMapping.cpp:30-58) of one position on the box
// [min, max), diam = max - min
__device__ inline void wrapInto(double4& p, const double* min, const double* max, const double* diam)
{
    double x[3] = {p.x, p.y, p.z};
    for (int d = 0; d < 3; ++d)
    {
        if (max[d] <= x[d])
        {
            x[d] -= diam[d];
            x[d] = fmax(x[d], min[d]);
        }
        if (x[d] < min[d])
        {
            x[d] += diam[d];
            if (max[d] <= x[d]) x[d] = min[d];
        }
    }
    p.x = x[0];
    p.y = x[1];
    p.z = x[2];
}
struct WrapBox
{
    double min[3], max[3], diam[3];
    const double4* from = nullptr;   ///< if not null: the current positions are from[slotOf[i]]
    const int32_t* slotOf = nullptr;
};

__device__ inline double coord(const double4& p, int axis)
{
    return axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
}

// a time step's post-force integration fused with the next step's pre-force integration
// (velocity Verlet, or Langevin without a predicate), bit-identical to the two calls;
// *maxDisp = sqrt(max |dx|^2) of the pre-force part (synchronises); sortedPos (if not
// null) receives each atom's new position at its sorted slot slotOf[i]
struct HistD;
// the fused post/pre step of the time loops; with adrMol (the AdResS loop, molecule k = atom
// k) each local atom's molecule force (ContributeMoleculeForceToAtoms) and the thermodynamic
// force of its x-bin (ThermodynamicForce::apply) are added to its force first.  With
// sortedPos / slotOf each new position is stored at the atom's slot too; slotsOnly: the
// positions are read from the slots and stored there alone (the positions array goes stale)
int fusedPostPre(mrmd_atoms* a, double dt, bool langevin, double zeta, double temperature, uint64_t seed,
                 uint64_t call, double* maxDisp, double4* sortedPos = nullptr, const int32_t* slotOf = nullptr,
                 const mrmd_molecules* adrMol = nullptr, const double* tfForce = nullptr,
                 const HistD* tfHd = nullptr, bool slotsOnly = false);
// whether mrmd_lj_apply without a predicate uses the full-row kernel (no ghost forces)
bool ljFullRows(const mrmd_lj* lj);
// device address of the allocation-free apply's scratch array on `device` (the current
// device); called at every context creation, looks it up once per device; false when the
// lookup failed
bool ljNoAllocSetup(int device);
const int32_t* ghostCorrDevice(const mrmd_ghost_layer* gl);
int64_t ghostCorrSize(const mrmd_ghost_layer* gl);
// fused passes of the AdResS time loop (molecule k = atom k; see adress.hip)
bool adressIdentityLayout(const mrmd_molecules* m, const mrmd_atoms* a);
void adressPrepare(mrmd_atoms* a, mrmd_molecules* m, const mrmd_ghost_layer* gl, const mrmd_box* box,
                   const mrmd_slab* slab, mrmd_lj_idealgas* ig, mrmd_verlet_list* vl);
void adressCollectGhosts(mrmd_atoms* a, mrmd_molecules* m, const mrmd_ghost_layer* gl);
void adressContributeLocal(mrmd_atoms* a, mrmd_molecules* m);
// the thermodynamic force's device table and histogram geometry
void thermoForceTable(const mrmd_thermo_force* tf, const double** force, HistD* hd);
// whether the table has the store form (single type, no capping): ljApplyAssign then
// computes what zeroing the forces of the list rows and mrmd_lj_apply(pred = NULL) leave
bool ljAssignable(const mrmd_lj* lj);
int ljApplyAssign(mrmd_lj* lj, mrmd_atoms* atoms, mrmd_verlet_list* vl);
// zero the forces of records [lo, hi)
int fillForceRange(mrmd_atoms* atoms, int64_t lo, int64_t hi);
// UpdateGhostAtoms::updateOnlyPos that also writes each ghost's new position to
// sortedPos[slotOf[ghost]] (the slot-order copy of a Verlet list); slotsOnly: every record's
// position is read from its slot and the ghost's stored there alone
int ghostUpdateSorted(mrmd_ghost_layer* gl, mrmd_atoms* a, const mrmd_box* box, double4* sortedPos,
                      const int32_t* slotOf, bool slotsOnly = false);
// PeriodicMapping of the real atoms whose current positions are in a list's slot-order copy:
// pos[i] = wrap(from[slotOf[i]]) for the local atoms
int haloWrapFromSlots(mrmd_atoms* a, const mrmd_box* box, const double4* from, const int32_t* slotOf);
}  // namespace mrmd_hip
