// The time loop of the reference's LJ examples, run natively over this library's entry
// points: examples/02_LennardJones_NVE/02_LennardJones_NVE.cpp:135-186 (pre-force
// integration, rebuild check, exchange + ghosts + list on a rebuild, ghost update
// otherwise, zero forces, LennardJones::apply, ghost force fold, post-force integration),
// with the Langevin pre-force integrator of examples/01_LennardJones_NVT.cpp:139-214 and the
// LinkedCellList + Atoms::permute of tests/NVT/NVT.cpp:138-144 on every rebuild when a cell
// list is given.  Each step makes the calls a caller of the actions makes, in the same
// order, except that a step's post-force integration and the next step's pre-force
// integration, adjacent in the loop, run as one pass over the atoms with the same
// arithmetic; the results are those of the same calls made one by one, bit for bit.
#include <cfloat>

#include "containers.hpp"
#include "nlist.hpp"
#include "physics.hpp"

using namespace mrmd_hip;

namespace
{
// positions kept in a Verlet list's slot-order copy alone (held) are copied back to the
// positions array by release() or, when a call leaves early, by the destructor
struct SlotPositions
{
    mrmd_verlet_list* vl;
    mrmd_atoms* a;
    bool held = false;
    void release()
    {
        if (held) vl->positionsFromSlots(a->pos.p);
        held = false;
    }
    ~SlotPositions()
    {
        if (held) vl->positionsFromSlotsNoThrow(a->pos.p);
    }
};
}  // namespace

extern "C" {

int mrmd_md_run(mrmd_atoms* a, mrmd_ghost_layer* gl, mrmd_verlet_list* vl, mrmd_lj* lj, mrmd_cell_list* cells,
                const mrmd_box* box, mrmd_md_loop* s, int64_t nsteps)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && gl && vl && lj && box && s && nsteps >= 0, "invalid arguments");
    REQUIRE(s->dt > 0.0 && s->neighbor_radius > 0.0 && s->skin >= 0.0 && s->cell_ratio > 0.0,
            "dt, neighbor_radius and cell_ratio must be positive, skin non-negative");
    const double rn = s->neighbor_radius;
    double delta[3];
    for (int d = 0; d < 3; ++d) delta[d] = rn;
    // slots.held: the current positions of every record are in the list's slot-order copy
    // alone (the positions array is stale); they are copied back before the call returns,
    // also when it returns with an error
    SlotPositions slots{vl, a};
#define MD_CALL(expr)                         \
    do                                        \
    {                                         \
        const int rc_ = (expr);               \
        if (rc_ != MRMD_OK) return rc_;       \
    } while (0)
    // a step's post-force integration and the next step's pre-force integration run as one
    // fused pass (fusedPostPre, bit-identical to the two calls); the first pre-force and the
    // last post-force integration of the call run alone
    double d = 0.0;
    if (nsteps > 0)
    {
        if (s->langevin)
            MD_CALL(mrmd_langevin_pre(a, s->dt, s->zeta, s->temperature, s->seed, s->call, nullptr, &d));
        else
            MD_CALL(mrmd_vv_pre(a, s->dt, &d));
    }
    // The force kernel reads positions in the list's sorted-slot order, normally gathered at
    // the start of every force call.  Here the passes that move the atoms write that copy
    // themselves: the fused post/pre pass stores each local atom's new position at its slot,
    // the ghost update each ghost's, and a rebuild's binning gathers it; the force call then
    // skips its gather (same values, so the same bits).  Between rebuilds the slot-order copy
    // is the only one kept: the post/pre pass and the ghost update read and store positions
    // at the slots alone, the rebuild's PeriodicMapping reads them from there into the
    // positions array, and the call's end copies them back (same values, one pass fewer of
    // position stores per step).
    bool localsScattered = false;  // the last pre-force pass wrote every local atom's slot
    // the ghost records' forces are zero: nothing between two rebuilds writes them on the
    // store-form path, so they are filled once after each ghost creation
    bool ghostForcesZero = false;
    for (int64_t k = 0; k < nsteps; ++k)
    {
        if (s->langevin) ++s->call;
        s->max_disp += d;
        bool fresh = false;
        if (s->max_disp >= s->skin * 0.5)
        {
            s->max_disp = 0.0;
            // the positions, from the slots when they are kept there, back in the positions array
            const double4* from = slots.held ? vl->bin.sortedPos.p : nullptr;
            const int32_t* slotOf = slots.held ? vl->slotOf.p : nullptr;
            // (the real atoms' positions are current in the positions array after the wrap; the
            // ghosts are made again from them)
            if (cells)
            {
                // PeriodicMapping and the cell list in one pass over the positions
                cellListBuildWrapped(cells, a, box, delta, from, slotOf);
                slots.held = false;
                // the ghosts are made again and every force computed again before either is read
                permuteAtoms(a, cells, false);
            }
            else if (from)
            {
                MD_CALL(haloWrapFromSlots(a, box, from, slotOf));
                slots.held = false;
            }
            else
                MD_CALL(mrmd_halo_wrap(a, box));
            int64_t ng = 0;
            MD_CALL(mrmd_halo_create(gl, a, box, &ng));
            ghostForcesZero = false;  // the new ghosts carry their atoms' forces
            MD_CALL(mrmd_verlet_build(vl, a, 0, a->nlocal, rn, s->cell_ratio, s->grid_min, s->grid_max,
                                      s->max_neigh));
            ++s->rebuilds;
            fresh = vl->nrows == a->nlocal + a->nghost;  // the build binned (and gathered) every record
        }
        else if (localsScattered && vl->nrows == a->nlocal + a->nghost)
        {
            MD_CALL(ghostUpdateSorted(gl, a, box, vl->bin.sortedPos.p, vl->ensureSlotOf(), slots.held));
            fresh = true;
        }
        else
        {
            slots.release();
            MD_CALL(mrmd_halo_update(gl, a, box));
        }
        vl->sortedFresh = fresh;
        if (ljAssignable(lj) && vl->begin == 0 && vl->end == a->nlocal)
        {
            // the force kernel stores the rows' forces (0.0 + F, what zeroing and adding
            // leave) and writes no ghost record; only the ghost records need their zero,
            // once per ghost creation.  A caller's list whose rows are not all local atoms
            // (possible before the loop's first rebuild) takes the fill-and-add path, so that
            // no local atom keeps an old force
            if (!ghostForcesZero) MD_CALL(fillForceRange(a, a->nlocal, a->nlocal + a->nghost));
            ghostForcesZero = true;
            MD_CALL(ljApplyAssign(lj, a, vl));
        }
        else
        {
            MD_CALL(mrmd_atoms_fill(a, MRMD_FORCE, 0.0));
            MD_CALL(mrmd_lj_apply(lj, a, vl, nullptr));
            ghostForcesZero = false;
        }
        // the full-row kernel leaves every ghost force at the zero just filled in, so the
        // fold (which adds only non-zero ghost forces) would change nothing
        if (!ljFullRows(lj)) MD_CALL(mrmd_halo_fold_forces(gl, a));
        localsScattered = false;
        if (k + 1 < nsteps)
        {
            const bool scatter = vl->ntiles > 0 && vl->nrows == a->nlocal + a->nghost;
            // the full-row force call has just read every record's position at its slot (its
            // own gather or the copies above), so the slots hold the current positions
            const bool slotsOnly = scatter && ljFullRows(lj);
            if (!slotsOnly) slots.release();
            MD_CALL(fusedPostPre(a, s->dt, s->langevin != 0, s->zeta, s->temperature, s->seed, s->call, &d,
                                 scatter ? vl->bin.sortedPos.p : nullptr, scatter ? vl->ensureSlotOf() : nullptr,
                                 nullptr, nullptr, nullptr, slotsOnly));
            localsScattered = scatter;
            if (slotsOnly) slots.held = true;
        }
        else
            MD_CALL(mrmd_vv_post(a, s->dt));
        ++s->steps;
    }
    slots.release();
#undef MD_CALL
    MRMD_API_END
}

// One step of the LJ / ideal-gas AdResS slab system in the order the actions' contracts
// force (molecule forces reach ghost atoms before the ghost fold; lambda is current before
// LJ_IdealGas::run): Langevin pre-force integration; on a rebuild UpdateMolecules (centres),
// MultiResPeriodicGhostExchange (exchange + ghosts), UpdateMolecules, the molecule Verlet
// list, otherwise a ghost refresh; UpdateMolecules with the slab weighting; zero atom and
// molecule forces; LJ_IdealGas::run; ContributeMoleculeForceToAtoms; ghost fold;
// ThermodynamicForce sample, update_if every tf_every steps, apply; post-force integration
// (action/LJ_IdealGas.cpp:227-260, action/ContributeMoleculeForceToAtoms.cpp:23-48,
// action/ThermodynamicForce.hpp:98-152).  The energy is read back on the last step only;
// post- and next pre-force integration run fused as in mrmd_md_run.
int mrmd_adress_md_run(mrmd_atoms* a, mrmd_molecules* m, mrmd_ghost_layer* gl, mrmd_verlet_list* vl,
                       mrmd_lj_idealgas* ig, mrmd_thermo_force* tf, const mrmd_slab* slab, const mrmd_box* box,
                       mrmd_adress_loop* s, int64_t nsteps)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && m && gl && vl && ig && tf && slab && box && s && nsteps >= 0, "invalid arguments");
    REQUIRE(s->dt > 0.0 && s->neighbor_radius > 0.0 && s->skin >= 0.0 && s->tf_every >= 0,
            "dt and neighbor_radius must be positive, skin and tf_every non-negative");
    const double rn = s->neighbor_radius;
    double gridMin[3], gridMax[3];
    for (int d = 0; d < 3; ++d)
    {
        gridMin[d] = box->min[d] - rn;
        gridMax[d] = box->max[d] + rn;
    }
#define MD_CALL(expr)                         \
    do                                        \
    {                                         \
        const int rc_ = (expr);               \
        if (rc_ != MRMD_OK) return rc_;       \
    } while (0)
    // post-force integration fused with the next step's pre-force integration, as above
    double d = 0.0;
    if (nsteps > 0) MD_CALL(mrmd_langevin_pre(a, s->dt, s->zeta, s->temperature, s->seed, s->call, nullptr, &d));
    // molecules of one atom each, molecule k = atom k (checked now and at every rebuild):
    // the per-record actions around LJ_IdealGas run as fused passes (adressPrepare before
    // it, adressCollectGhosts after it, the local molecule forces and the thermodynamic force
    // inside the post/pre step)
    bool fused = nsteps > 0 && adressIdentityLayout(m, a);
    const double* tfForce = nullptr;
    HistD tfHd{};
    thermoForceTable(tf, &tfForce, &tfHd);
    for (int64_t k = 0; k < nsteps; ++k)
    {
        ++s->call;
        s->max_disp += d;
        const bool rebuild = s->max_disp >= s->skin * 0.5;
        if (rebuild)
        {
            s->max_disp = 0.0;
            MD_CALL(mrmd_molecules_update(m, a, nullptr));
            MD_CALL(mrmd_multires_wrap(m, a, box));
            MD_CALL(mrmd_multires_halo_create(gl, m, a, box));
            MD_CALL(mrmd_molecules_update(m, a, nullptr));
            // monatomic molecules (the fused layout): rows only where a pair can carry a weight
            // before the next rebuild, the x band of the slab's atomistic and hybrid regions
            // widened by the displacement that triggers a rebuild (a molecule's move since
            // the build stays below skin / 2); a layout that is no longer monatomic builds
            // again in full
            if (fused && slab)
            {
                const double reach = 0.5 * slab->atomistic_diameter + slab->hybrid_diameter + 0.5 * s->skin;
                vl->bandLo = slab->center[0] - reach;
                vl->bandHi = slab->center[0] + reach;
            }
            const int built = mrmd_verlet_build_molecules(vl, m, 0, m->nlocal, rn, 1.0, gridMin, gridMax, 0);
            vl->bandLo = 1.0;
            vl->bandHi = -1.0;
            MD_CALL(built);
            ++s->rebuilds;
            const bool wasBanded = vl->banded;
            fused = adressIdentityLayout(m, a);
            if (wasBanded && !fused)
                MD_CALL(mrmd_verlet_build_molecules(vl, m, 0, m->nlocal, rn, 1.0, gridMin, gridMax, 0));
        }
        if (fused)
            adressPrepare(a, m, rebuild ? nullptr : gl, box, slab, ig, vl);
        else
        {
            if (!rebuild) MD_CALL(mrmd_halo_update(gl, a, box));
            MD_CALL(mrmd_molecules_update(m, a, slab));
            MD_CALL(mrmd_atoms_fill(a, MRMD_FORCE, 0.0));
            MD_CALL(mrmd_molecules_fill(m, MRMD_MOL_FORCE, 0.0));
        }
        MD_CALL(mrmd_ljig_run(ig, m, vl, a, k + 1 == nsteps ? &s->energy : nullptr));
        if (fused)
            adressCollectGhosts(a, m, gl);
        else
        {
            MD_CALL(mrmd_molecules_contribute_force(m, a));
            MD_CALL(mrmd_halo_fold_forces(gl, a));
        }
        MD_CALL(mrmd_tf_sample(tf, a));
        ++s->steps;
        if (s->tf_every && s->steps % s->tf_every == 0)
            MD_CALL(mrmd_tf_update(tf, s->tf_sigma, s->tf_range, &s->tf_pred));
        if (k + 1 < nsteps && fused)
            MD_CALL(fusedPostPre(a, s->dt, true, s->zeta, s->temperature, s->seed, s->call, &d, nullptr, nullptr, m,
                                 tfForce, &tfHd));
        else
        {
            if (fused) adressContributeLocal(a, m);
            MD_CALL(mrmd_tf_apply(tf, a, nullptr));
            if (k + 1 < nsteps)
                MD_CALL(fusedPostPre(a, s->dt, true, s->zeta, s->temperature, s->seed, s->call, &d));
            else
                MD_CALL(mrmd_vv_post(a, s->dt));
        }
    }
#undef MD_CALL
    MRMD_API_END
}

}  // extern "C"
