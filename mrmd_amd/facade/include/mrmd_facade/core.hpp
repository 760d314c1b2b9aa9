// C++ host API of the gfx950 library with the reference's class names and signatures
// (namespace mrmd: data::Atoms, communication::GhostLayer, HalfVerletList,
// action::LennardJones, ...).  Every call goes through the C ABI of mrmd_hip.h; device
// memory belongs to the library.  Differences to the reference are the ones a C ABI
// forces: slices are handles (read them on the host through HostAtoms/deep_copy), and
// predicates are descriptor functors (pred::*, util::IsInSymmetricSlab, region functors)
// rather than arbitrary device lambdas.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "mrmd_hip.h"

namespace mrmd
{
// ---- core types (mrmd/datatypes.hpp:91-151, constants.hpp:24-27) -----------------------
using idx_t = int64_t;
using real_t = double;
using Point3D = std::array<real_t, 3>;
using Vector3D = std::array<real_t, 3>;
constexpr real_t operator""_r(long double v) { return static_cast<real_t>(v); }
constexpr real_t operator""_r(unsigned long long v) { return static_cast<real_t>(v); }
template <class T>
constexpr idx_t idx_c(const T& v)
{
    return static_cast<idx_t>(v);
}
template <class T>
constexpr real_t real_c(T v)
{
    return static_cast<real_t>(v);
}
constexpr int DIMENSIONS = 3;
constexpr real_t pi = 3.14159265358979323846;
enum class AXIS : idx_t
{
    X = 0,
    Y = 1,
    Z = 2
};
template <class E>
constexpr auto to_underlying(E e) noexcept
{
    return static_cast<std::underlying_type_t<E>>(e);
}

// ---- context -----------------------------------------------------------------------------
namespace impl
{
inline mrmd_ctx*& ctxSlot()
{
    static mrmd_ctx* c = nullptr;
    return c;
}
inline void check(int rc, const char* what)
{
    if (rc != MRMD_OK)
    {
        const char* msg = ctxSlot() ? mrmd_ctx_last_error(ctxSlot()) : "";
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + msg);
    }
}
}  // namespace impl

/// replaces Kokkos::ScopeGuard / mrmd::initialize (initialization.cpp:23-44)
inline void initialize(int device = 0)
{
    if (!impl::ctxSlot()) impl::check(mrmd_ctx_create(device, &impl::ctxSlot()), "mrmd_ctx_create");
}
inline void finalize()
{
    if (impl::ctxSlot()) mrmd_ctx_destroy(impl::ctxSlot());
    impl::ctxSlot() = nullptr;
}
inline mrmd_ctx* context()
{
    initialize();
    return impl::ctxSlot();
}
struct ScopeGuard
{
    explicit ScopeGuard(int device = 0) { initialize(device); }
    ScopeGuard(int, char**) { initialize(0); }
    ~ScopeGuard() { finalize(); }
};
inline void fence() { impl::check(mrmd_ctx_synchronize(context()), "fence"); }

// ---- predicates --------------------------------------------------------------------------
// Descriptor functors: each maps to mrmd_pred1 (and evaluates the same on the host).
namespace pred
{
struct Always
{
    mrmd_pred1 desc() const { return mrmd_pred1{MRMD_PRED_ALWAYS, 0, 0, 0, 0, 0}; }
};
struct Never
{
    mrmd_pred1 desc() const { return mrmd_pred1{MRMD_PRED_NEVER, 0, 0, 0, 0, 0}; }
};
/// lower < coordinate(axis) < upper
struct OpenRange
{
    real_t lower, upper;
    AXIS axis = AXIS::X;
    mrmd_pred1 desc() const
    {
        return mrmd_pred1{MRMD_PRED_OPEN_RANGE, int32_t(to_underlying(axis)), lower, upper, 0, 0};
    }
};
struct Greater
{
    real_t value;
    AXIS axis = AXIS::X;
    mrmd_pred1 desc() const { return mrmd_pred1{MRMD_PRED_GREATER, int32_t(to_underlying(axis)), value, 0, 0, 0}; }
};
struct Less
{
    real_t value;
    AXIS axis = AXIS::X;
    mrmd_pred1 desc() const { return mrmd_pred1{MRMD_PRED_LESS, int32_t(to_underlying(axis)), value, 0, 0, 0}; }
};
/// op(first(pos_i), second(pos_j)) for TwoPositionsPredicate call sites
template <class A, class B>
struct Or
{
    A first;
    B second;
    mrmd_pred2 desc2() const { return mrmd_pred2{MRMD_PAIR_OR, 0, first.desc(), second.desc()}; }
};
template <class A, class B>
struct And
{
    A first;
    B second;
    mrmd_pred2 desc2() const { return mrmd_pred2{MRMD_PAIR_AND, 0, first.desc(), second.desc()}; }
};
template <class A, class B>
Or<A, B> either(A a, B b)
{
    return {a, b};
}
template <class A, class B>
And<A, B> both(A a, B b)
{
    return {a, b};
}
}  // namespace pred

namespace util
{
/// util/IsInSymmetricSlab.hpp:24-64
class IsInSymmetricSlab
{
    real_t center_, slabMin_, slabMax_;
    AXIS axis_;
    real_t tolerance_;

public:
    IsInSymmetricSlab(const Point3D& center, real_t slabMin, real_t slabMax, AXIS axis = AXIS::X,
                      real_t tolerance = 0_r)
        : center_(center[to_underlying(axis)]), slabMin_(slabMin), slabMax_(slabMax), axis_(axis), tolerance_(tolerance)
    {
    }
    bool operator()(real_t x, real_t y, real_t z) const
    {
        const real_t c[3] = {x, y, z};
        return (*this)(c[to_underlying(axis_)]);
    }
    bool operator()(real_t coord) const
    {
        auto absDx = std::abs(coord - center_);
        return absDx >= slabMin_ - tolerance_ && absDx <= slabMax_ + tolerance_;
    }
    mrmd_pred1 desc() const
    {
        return mrmd_pred1{MRMD_PRED_SYMMETRIC_SLAB, int32_t(to_underlying(axis_)), center_, slabMin_, slabMax_,
                          tolerance_};
    }
};

/// util/math.hpp:31-46
inline real_t powInt(const real_t& x, const idx_t n)
{
    real_t ww = x, yy = 1_r;
    for (idx_t nn = (n > 0) ? n : -n; nn != 0; nn >>= 1)
    {
        if ((nn & 1) == 1) yy *= ww;
        ww *= ww;
    }
    return (n > 0) ? yy : 1.0 / yy;
}
inline real_t sqr(const real_t& x) { return x * x; }
/// util/interpolation.hpp:34-38
inline real_t lerp(const real_t& left, const real_t& right, const real_t& factor)
{
    return left + (right - left) * factor;
}
}  // namespace util

template <class P>
inline mrmd_pred1 toPred1(const P& p)
{
    return p.desc();
}
template <class P, class = void>
struct HasDesc2 : std::false_type
{
};
template <class P>
struct HasDesc2<P, std::void_t<decltype(std::declval<P>().desc2())>> : std::true_type
{
};

// ---- data ------------------------------------------------------------------------------
namespace data
{
/// data::Subdomain (mrmd/data/Subdomain.hpp:37-110)
struct Subdomain
{
    Subdomain() = default;
    Subdomain(const Point3D& minC, const Point3D& maxC, const Vector3D& ghost)
        : minCorner(minC), maxCorner(maxC), ghostLayerThickness(ghost)
    {
        for (int d = 0; d < 3; ++d)
        {
            if (ghost[d] < 0) throw std::invalid_argument("negative ghost layer");
            minGhostCorner[d] = minCorner[d] - ghostLayerThickness[d];
            maxGhostCorner[d] = maxCorner[d] + ghostLayerThickness[d];
            minInnerCorner[d] = minCorner[d] + ghostLayerThickness[d];
            maxInnerCorner[d] = maxCorner[d] - ghostLayerThickness[d];
            diameter[d] = maxCorner[d] - minCorner[d];
            diameterWithGhostLayer[d] = maxCorner[d] - minCorner[d] + 2_r * ghostLayerThickness[d];
        }
    }
    Subdomain(const Point3D& minC, const Point3D& maxC, const real_t& ghost) : Subdomain(minC, maxC, {ghost, ghost, ghost})
    {
    }
    real_t getVolume() const { return diameter[0] * diameter[1] * diameter[2]; }
    Point3D getCenter() const
    {
        return {(minCorner[0] + maxCorner[0]) * 0.5_r, (minCorner[1] + maxCorner[1]) * 0.5_r,
                (minCorner[2] + maxCorner[2]) * 0.5_r};
    }
    mrmd_box box() const
    {
        mrmd_box b;
        for (int d = 0; d < 3; ++d)
        {
            b.min[d] = minCorner[d];
            b.max[d] = maxCorner[d];
            b.ghost[d] = ghostLayerThickness[d];
        }
        return b;
    }
    Point3D minCorner{}, maxCorner{}, ghostLayerThickness{}, minGhostCorner{}, maxGhostCorner{}, minInnerCorner{},
        maxInnerCorner{};
    Vector3D diameter{}, diameterWithGhostLayer{};
};

/// host-side record view used by HostAtoms / HostMolecules slices: s(i, d) and s(i)
template <class T, int W>
struct HostSlice
{
    std::vector<T>* v = nullptr;
    T& operator()(idx_t i, int d = 0) const { return (*v)[size_t(i * W + d)]; }
    idx_t size() const { return idx_t(v->size() / W); }
};

class Atoms;

/// data::HostAtoms: host arrays with the reference's slice accessors
class HostAtoms
{
public:
    explicit HostAtoms(idx_t n = 0) { resize(n); }
    explicit HostAtoms(const Atoms& a);
    void resize(size_t n)
    {
        pos_.resize(3 * n, 0_r);
        vel_.resize(3 * n, 0_r);
        force_.resize(3 * n, 0_r);
        type_.resize(n, 0);
        mass_.resize(n, 0_r);
        charge_.resize(n, 0_r);
        relMass_.resize(n, 0_r);
    }
    idx_t size() const { return idx_t(type_.size()); }
    HostSlice<real_t, 3> getPos() { return {&pos_}; }
    HostSlice<real_t, 3> getVel() { return {&vel_}; }
    HostSlice<real_t, 3> getForce() { return {&force_}; }
    HostSlice<idx_t, 1> getType() { return {&type_}; }
    HostSlice<real_t, 1> getMass() { return {&mass_}; }
    HostSlice<real_t, 1> getCharge() { return {&charge_}; }
    HostSlice<real_t, 1> getRelativeMass() { return {&relMass_}; }
    idx_t numLocalAtoms = 0;
    idx_t numGhostAtoms = 0;

    std::vector<real_t> pos_, vel_, force_, mass_, charge_, relMass_;
    std::vector<idx_t> type_;
};

class Molecules;

/// handle to one field of a device container (what getPos() etc. return)
struct DeviceSlice
{
    const Atoms* atoms;
    int field;
    const Molecules* molecules = nullptr;
};

/// data::Atoms (mrmd/data/Atoms.hpp:33-177) in device memory
class Atoms
{
public:
    explicit Atoms(idx_t numAtoms)
    {
        impl::check(mrmd_atoms_create(context(), numAtoms, &h_), "mrmd_atoms_create");
    }
    Atoms(const Atoms&) = delete;
    Atoms& operator=(const Atoms&) = delete;
    Atoms(Atoms&& o) noexcept { *this = std::move(o); }
    Atoms& operator=(Atoms&& o) noexcept
    {
        std::swap(h_, o.h_);
        numLocalAtoms = o.numLocalAtoms;
        numGhostAtoms = o.numGhostAtoms;
        return *this;
    }
    explicit Atoms(const HostAtoms& host) : Atoms(host.size()) { copyFrom(host); }
    ~Atoms()
    {
        if (h_) mrmd_atoms_destroy(h_);
    }

    idx_t size() const
    {
        int64_t nl, ng, sz;
        mrmd_atoms_get_counts(h_, &nl, &ng, &sz);
        return sz;
    }
    void resize(size_t n) { impl::check(mrmd_atoms_resize(h_, int64_t(n)), "Atoms::resize"); }
    void setForce(const real_t& v) const { impl::check(mrmd_atoms_fill(h_, MRMD_FORCE, v), "Atoms::setForce"); }
    void removeGhostAtoms()
    {
        numGhostAtoms = 0;
        resize(size_t(numLocalAtoms));
    }
    DeviceSlice getPos() const { return {this, MRMD_POS}; }
    DeviceSlice getVel() const { return {this, MRMD_VEL}; }
    DeviceSlice getForce() const { return {this, MRMD_FORCE}; }
    DeviceSlice getType() const { return {this, MRMD_TYPE}; }
    DeviceSlice getMass() const { return {this, MRMD_MASS}; }
    DeviceSlice getCharge() const { return {this, MRMD_CHARGE}; }
    DeviceSlice getRelativeMass() const { return {this, MRMD_RELMASS}; }

    /// push the public counters to the library (done before every action)
    mrmd_atoms* sync() const
    {
        impl::check(mrmd_atoms_set_counts(h_, numLocalAtoms, numGhostAtoms), "Atoms counts");
        return h_;
    }
    /// pull the counters after an action that changes them
    void pull()
    {
        int64_t nl, ng, sz;
        mrmd_atoms_get_counts(h_, &nl, &ng, &sz);
        numLocalAtoms = nl;
        numGhostAtoms = ng;
    }
    void copyFrom(const HostAtoms& host)
    {
        const int64_t n = host.size();
        if (n > size()) resize(size_t(n));
        impl::check(mrmd_atoms_upload(h_, MRMD_POS, host.pos_.data(), 0, n), "upload");
        impl::check(mrmd_atoms_upload(h_, MRMD_VEL, host.vel_.data(), 0, n), "upload");
        impl::check(mrmd_atoms_upload(h_, MRMD_FORCE, host.force_.data(), 0, n), "upload");
        impl::check(mrmd_atoms_upload(h_, MRMD_TYPE, host.type_.data(), 0, n), "upload");
        impl::check(mrmd_atoms_upload(h_, MRMD_MASS, host.mass_.data(), 0, n), "upload");
        impl::check(mrmd_atoms_upload(h_, MRMD_CHARGE, host.charge_.data(), 0, n), "upload");
        impl::check(mrmd_atoms_upload(h_, MRMD_RELMASS, host.relMass_.data(), 0, n), "upload");
        numLocalAtoms = host.numLocalAtoms;
        numGhostAtoms = host.numGhostAtoms;
        sync();
    }
    void copyTo(HostAtoms& host) const
    {
        const int64_t n = size();
        host.resize(size_t(n));
        if (n)
        {
            impl::check(mrmd_atoms_download(h_, MRMD_POS, host.pos_.data(), 0, n), "download");
            impl::check(mrmd_atoms_download(h_, MRMD_VEL, host.vel_.data(), 0, n), "download");
            impl::check(mrmd_atoms_download(h_, MRMD_FORCE, host.force_.data(), 0, n), "download");
            impl::check(mrmd_atoms_download(h_, MRMD_TYPE, host.type_.data(), 0, n), "download");
            impl::check(mrmd_atoms_download(h_, MRMD_MASS, host.mass_.data(), 0, n), "download");
            impl::check(mrmd_atoms_download(h_, MRMD_CHARGE, host.charge_.data(), 0, n), "download");
            impl::check(mrmd_atoms_download(h_, MRMD_RELMASS, host.relMass_.data(), 0, n), "download");
        }
        host.numLocalAtoms = numLocalAtoms;
        host.numGhostAtoms = numGhostAtoms;
    }
    mrmd_atoms* handle() const { return h_; }

    idx_t numLocalAtoms = 0;
    idx_t numGhostAtoms = 0;

    template <class CellList>
    void permute(const CellList& lcl)
    {
        impl::check(mrmd_atoms_permute(sync(), lcl.handle()), "Atoms::permute");
    }

private:
    mrmd_atoms* h_ = nullptr;
};

inline HostAtoms::HostAtoms(const Atoms& a) { a.copyTo(*this); }
inline void deep_copy(HostAtoms& dst, const Atoms& src) { src.copyTo(dst); }
inline void deep_copy(Atoms& dst, const HostAtoms& src) { dst.copyFrom(src); }

class Molecules;

/// data::HostMolecules
class HostMolecules
{
public:
    explicit HostMolecules(idx_t n = 0) { resize(size_t(n)); }
    explicit HostMolecules(const Molecules& m);
    void resize(size_t n)
    {
        pos_.resize(3 * n, 0_r);
        force_.resize(3 * n, 0_r);
        grad_.resize(3 * n, 0_r);
        lambda_.resize(n, 0_r);
        modLambda_.resize(n, 0_r);
        offset_.resize(n, 0);
        numAtoms_.resize(n, 0);
    }
    idx_t size() const { return idx_t(offset_.size()); }
    HostSlice<real_t, 3> getPos() { return {&pos_}; }
    HostSlice<real_t, 3> getForce() { return {&force_}; }
    HostSlice<real_t, 3> getGradLambda() { return {&grad_}; }
    HostSlice<real_t, 1> getLambda() { return {&lambda_}; }
    HostSlice<real_t, 1> getModulatedLambda() { return {&modLambda_}; }
    HostSlice<idx_t, 1> getAtomsOffset() { return {&offset_}; }
    HostSlice<idx_t, 1> getNumAtoms() { return {&numAtoms_}; }
    idx_t numLocalMolecules = 0;
    idx_t numGhostMolecules = 0;
    std::vector<real_t> pos_, force_, grad_, lambda_, modLambda_;
    std::vector<idx_t> offset_, numAtoms_;
};

/// data::Molecules (mrmd/data/Molecules.hpp:27-176) in device memory
class Molecules
{
public:
    explicit Molecules(idx_t n) { impl::check(mrmd_molecules_create(context(), n, &h_), "mrmd_molecules_create"); }
    Molecules(const Molecules&) = delete;
    Molecules(Molecules&& o) noexcept { *this = std::move(o); }
    Molecules& operator=(Molecules&& o) noexcept
    {
        std::swap(h_, o.h_);
        numLocalMolecules = o.numLocalMolecules;
        numGhostMolecules = o.numGhostMolecules;
        return *this;
    }
    explicit Molecules(const HostMolecules& host) : Molecules(host.size()) { copyFrom(host); }
    ~Molecules()
    {
        if (h_) mrmd_molecules_destroy(h_);
    }
    idx_t size() const
    {
        int64_t nl, ng, sz;
        mrmd_molecules_get_counts(h_, &nl, &ng, &sz);
        return sz;
    }
    void resize(size_t n) { impl::check(mrmd_molecules_resize(h_, int64_t(n)), "Molecules::resize"); }
    void setForce(const real_t& v) const { impl::check(mrmd_molecules_fill(h_, MRMD_MOL_FORCE, v), "setForce"); }
    DeviceSlice getPos() const { return {nullptr, MRMD_MOL_POS, this}; }
    mrmd_molecules* sync() const
    {
        impl::check(mrmd_molecules_set_counts(h_, numLocalMolecules, numGhostMolecules), "Molecules counts");
        return h_;
    }
    void pull()
    {
        int64_t nl, ng, sz;
        mrmd_molecules_get_counts(h_, &nl, &ng, &sz);
        numLocalMolecules = nl;
        numGhostMolecules = ng;
    }
    void copyFrom(const HostMolecules& host)
    {
        const int64_t n = host.size();
        if (n > size()) resize(size_t(n));
        impl::check(mrmd_molecules_upload(h_, MRMD_MOL_POS, host.pos_.data(), 0, n), "upload");
        impl::check(mrmd_molecules_upload(h_, MRMD_MOL_FORCE, host.force_.data(), 0, n), "upload");
        impl::check(mrmd_molecules_upload(h_, MRMD_MOL_GRADLAMBDA, host.grad_.data(), 0, n), "upload");
        impl::check(mrmd_molecules_upload(h_, MRMD_MOL_LAMBDA, host.lambda_.data(), 0, n), "upload");
        impl::check(mrmd_molecules_upload(h_, MRMD_MOL_MODLAMBDA, host.modLambda_.data(), 0, n), "upload");
        impl::check(mrmd_molecules_upload(h_, MRMD_MOL_OFFSET, host.offset_.data(), 0, n), "upload");
        impl::check(mrmd_molecules_upload(h_, MRMD_MOL_NUMATOMS, host.numAtoms_.data(), 0, n), "upload");
        numLocalMolecules = host.numLocalMolecules;
        numGhostMolecules = host.numGhostMolecules;
        sync();
    }
    void copyTo(HostMolecules& host) const
    {
        const int64_t n = size();
        host.resize(size_t(n));
        if (n)
        {
            impl::check(mrmd_molecules_download(h_, MRMD_MOL_POS, host.pos_.data(), 0, n), "download");
            impl::check(mrmd_molecules_download(h_, MRMD_MOL_FORCE, host.force_.data(), 0, n), "download");
            impl::check(mrmd_molecules_download(h_, MRMD_MOL_GRADLAMBDA, host.grad_.data(), 0, n), "download");
            impl::check(mrmd_molecules_download(h_, MRMD_MOL_LAMBDA, host.lambda_.data(), 0, n), "download");
            impl::check(mrmd_molecules_download(h_, MRMD_MOL_MODLAMBDA, host.modLambda_.data(), 0, n), "download");
            impl::check(mrmd_molecules_download(h_, MRMD_MOL_OFFSET, host.offset_.data(), 0, n), "download");
            impl::check(mrmd_molecules_download(h_, MRMD_MOL_NUMATOMS, host.numAtoms_.data(), 0, n), "download");
        }
        host.numLocalMolecules = numLocalMolecules;
        host.numGhostMolecules = numGhostMolecules;
    }
    mrmd_molecules* handle() const { return h_; }
    idx_t numLocalMolecules = 0;
    idx_t numGhostMolecules = 0;

private:
    mrmd_molecules* h_ = nullptr;
};

inline HostMolecules::HostMolecules(const Molecules& m) { m.copyTo(*this); }
inline void deep_copy(HostMolecules& dst, const Molecules& src) { src.copyTo(dst); }
inline void deep_copy(Molecules& dst, const HostMolecules& src) { dst.copyFrom(src); }

/// data::createMoleculeForEachAtom (mrmd/data/MoleculesFromAtoms.cpp:19-39)
inline Molecules createMoleculeForEachAtom(Atoms& atoms)
{
    Molecules m(0);
    impl::check(mrmd_molecules_for_each_atom(m.handle(), atoms.sync()), "createMoleculeForEachAtom");
    m.pull();
    return m;
}

/// data::MultiHistogram (mrmd/data/MultiHistogram.hpp:29-92), host resident
struct MultiHistogram
{
    MultiHistogram(const std::string&, real_t minArg, real_t maxArg, idx_t numBinsArg, idx_t numHistogramsArg)
        : min(minArg),
          max(maxArg),
          numBins(numBinsArg),
          numHistograms(numHistogramsArg),
          binSize((maxArg - minArg) / real_c(numBinsArg)),
          inverseBinSize(1_r / binSize),
          data(size_t(numBinsArg * numHistogramsArg), 0_r)
    {
    }
    idx_t getBin(const real_t& val) const
    {
        auto bin = idx_c(std::floor((val - min) * inverseBinSize));
        if (bin < 0) bin = -1;
        if (bin >= numBins) bin = -1;
        return bin;
    }
    real_t getBinPosition(idx_t b) const { return min + (real_c(b) + 0.5_r) * binSize; }
    real_t& operator()(idx_t bin, idx_t hist) { return data[size_t(bin * numHistograms + hist)]; }
    real_t operator()(idx_t bin, idx_t hist) const { return data[size_t(bin * numHistograms + hist)]; }
    real_t min, max;
    idx_t numBins, numHistograms;
    real_t binSize, inverseBinSize;
    std::vector<real_t> data;
};
}  // namespace data

// ---- neighbour search -------------------------------------------------------------------
/// Cabana::LinkedCellList(pos, begin, end, delta, min, max) with stable in-cell order
class LinkedCellList
{
public:
    LinkedCellList(const data::DeviceSlice& pos, idx_t begin, idx_t end, const real_t* delta, const real_t* gridMin,
                   const real_t* gridMax)
    {
        impl::check(mrmd_cell_list_create(context(), &h_), "mrmd_cell_list_create");
        impl::check(mrmd_cell_list_build(h_, pos.atoms->sync(), begin, end, delta, gridMin, gridMax),
                    "LinkedCellList build");
    }
    LinkedCellList(const LinkedCellList&) = delete;
    ~LinkedCellList() { mrmd_cell_list_destroy(h_); }
    mrmd_cell_list* handle() const { return h_; }

private:
    mrmd_cell_list* h_ = nullptr;
};

/// HalfVerletList = Cabana::VerletList<HalfNeighborTag, VerletLayout2D, TeamOpTag>
class HalfVerletList
{
public:
    HalfVerletList() = default;
    HalfVerletList(const data::DeviceSlice& pos, idx_t begin, idx_t end, real_t r, real_t cellRatio,
                   const real_t* gridMin, const real_t* gridMax, idx_t maxNeigh = 0)
    {
        build(pos, begin, end, r, cellRatio, gridMin, gridMax, maxNeigh);
    }
    HalfVerletList(const HalfVerletList&) = delete;
    ~HalfVerletList()
    {
        if (h_) mrmd_verlet_destroy(h_);
    }
    void build(const data::DeviceSlice& pos, idx_t begin, idx_t end, real_t r, real_t cellRatio, const real_t* gridMin,
               const real_t* gridMax, idx_t maxNeigh = 0)
    {
        ensure();
        if (pos.atoms)
            impl::check(mrmd_verlet_build(h_, pos.atoms->sync(), begin, end, r, cellRatio, gridMin, gridMax, maxNeigh),
                        "HalfVerletList::build");
        else if (pos.molecules)
            buildMolecules(*pos.molecules, begin, end, r, cellRatio, gridMin, gridMax, maxNeigh);
        else
            throw std::invalid_argument("HalfVerletList::build needs a position slice");
        host_.clear();
    }
    /// the molecule-centre list of LJ_IdealGas (LJ_IdealGas.test.cpp:108-120)
    void buildMolecules(const data::Molecules& m, idx_t begin, idx_t end, real_t r, real_t cellRatio,
                        const real_t* gridMin, const real_t* gridMax, idx_t maxNeigh = 0)
    {
        ensure();
        impl::check(mrmd_verlet_build_molecules(h_, m.sync(), begin, end, r, cellRatio, gridMin, gridMax, maxNeigh),
                    "HalfVerletList::build (molecules)");
        host_.clear();
    }
    /// total pairs of the half list (sum of counts)
    idx_t numPairs() const
    {
        int64_t t = 0;
        impl::check(mrmd_verlet_half_counts(h_, nullptr, 0, &t), "numPairs");
        return t;
    }
    mrmd_verlet_list* handle() const { return h_; }

    /// host copy of the half list for HalfNeighborList::numNeighbor / getNeighbor
    void fetch(idx_t n) const
    {
        counts_.assign(size_t(n), 0);
        int64_t t = 0;
        impl::check(mrmd_verlet_half_counts(h_, counts_.data(), n, &t), "half counts");
        width_ = 1;
        for (auto c : counts_) width_ = std::max<idx_t>(width_, c);
        host_.assign(size_t(n * width_), -1);
        impl::check(mrmd_verlet_half_neighbors(h_, host_.data(), n, width_), "half neighbors");
    }
    mutable std::vector<int32_t> counts_, host_;
    mutable idx_t width_ = 1;

private:
    void ensure()
    {
        if (!h_) impl::check(mrmd_verlet_create(context(), &h_), "mrmd_verlet_create");
    }
    mrmd_verlet_list* h_ = nullptr;
};

/// Cabana::NeighborList<HalfVerletList> host accessors (after HalfVerletList::fetch)
struct HalfNeighborList
{
    static idx_t numNeighbor(const HalfVerletList& l, idx_t i) { return l.counts_[size_t(i)]; }
    static idx_t getNeighbor(const HalfVerletList& l, idx_t i, idx_t n) { return l.host_[size_t(i * l.width_ + n)]; }
};

// ---- weighting ---------------------------------------------------------------------------
namespace weighting_function
{
/// CheckRegion.hpp:25-38
constexpr real_t REGION_CHECK_EPSILON = 0_r;
inline bool isInATRegion(const real_t& l) { return l >= (1_r - REGION_CHECK_EPSILON); }
inline bool isInCGRegion(const real_t& l) { return l <= REGION_CHECK_EPSILON; }
inline bool isInHYRegion(const real_t& l) { return !isInATRegion(l) && !isInCGRegion(l); }

/// weighting_function::Slab (Slab.hpp:27-201)
class Slab
{
public:
    enum class InterfaceType
    {
        SMOOTH,
        ABRUPT
    };
    Slab(const Point3D& center, real_t atomisticRegionDiameter, real_t hybridRegionDiameter, idx_t nu,
         InterfaceType it = InterfaceType::SMOOTH)
    {
        for (int d = 0; d < 3; ++d) p_.center[d] = center[d];
        p_.atomistic_diameter = atomisticRegionDiameter;
        p_.hybrid_diameter = hybridRegionDiameter;
        p_.nu = nu;
        p_.interface_type = it == InterfaceType::ABRUPT ? 1 : 0;
        p_.pad = 0;
    }
    bool isInATRegion(real_t x, real_t, real_t) const { return std::abs(x - p_.center[0]) < half(); }
    bool isInHYRegion(real_t x, real_t, real_t) const
    {
        auto a = std::abs(x - p_.center[0]);
        if (a < half()) return false;
        return a < half() + p_.hybrid_diameter;
    }
    bool isInCGRegion(real_t x, real_t, real_t) const
    {
        return !(std::abs(x - p_.center[0]) < half() + p_.hybrid_diameter);
    }
    /// operator() evaluated by the device library (bit-identical to the kernels)
    void operator()(real_t x, real_t, real_t, real_t& lambda, real_t& modLambda, real_t& gx, real_t& gy,
                    real_t& gz) const
    {
        real_t out[5];
        impl::check(mrmd_slab_eval(context(), &p_, &x, 1, out), "Slab");
        lambda = out[0];
        modLambda = out[1];
        gx = out[2];
        gy = out[3];
        gz = out[4];
    }
    const mrmd_slab& params() const { return p_; }
    /// region predicates as descriptors (for apply_if / preForceIntegrate_apply_if)
    mrmd_pred1 atRegion() const { return {MRMD_PRED_SLAB_AT, 0, p_.center[0], half(), p_.hybrid_diameter, 0}; }
    mrmd_pred1 hyRegion() const { return {MRMD_PRED_SLAB_HY, 0, p_.center[0], half(), p_.hybrid_diameter, 0}; }
    mrmd_pred1 cgRegion() const { return {MRMD_PRED_SLAB_CG, 0, p_.center[0], half(), p_.hybrid_diameter, 0}; }

private:
    real_t half() const { return 0.5_r * p_.atomistic_diameter; }
    mrmd_slab p_{};
};
}  // namespace weighting_function

namespace weighting_function
{
/// weighting_function::Spherical (Spherical.hpp:25-99), evaluated by the device library
class Spherical
{
public:
    Spherical(const Point3D& center, const real_t atomisticRadius, const real_t hybridRegionDiameter,
              const int exponent)
    {
        for (int d = 0; d < 3; ++d) p_.center[d] = center[d];
        p_.atomistic_radius = atomisticRadius;
        p_.hybrid_diameter = hybridRegionDiameter;
        p_.exponent = exponent;
    }
    void operator()(const real_t x, const real_t y, const real_t z, real_t& lambda, real_t& gradLambdaX,
                    real_t& gradLambdaY, real_t& gradLambdaZ) const
    {
        const real_t p[3] = {x, y, z};
        real_t out[4];
        impl::check(mrmd_spherical_eval(context(), &p_, p, 1, out), "Spherical");
        lambda = out[0];
        gradLambdaX = out[1];
        gradLambdaY = out[2];
        gradLambdaZ = out[3];
    }
    real_t operator()(const real_t x, const real_t y, const real_t z) const
    {
        real_t lambda, tmp;
        operator()(x, y, z, lambda, tmp, tmp, tmp);
        return lambda;
    }
    const mrmd_spherical& params() const { return p_; }

private:
    mrmd_spherical p_{};
};
}  // namespace weighting_function

namespace pred
{
/// region of a slab as a one-position predicate
struct SlabRegion
{
    mrmd_pred1 d;
    mrmd_pred1 desc() const { return d; }
};
inline SlabRegion inAT(const weighting_function::Slab& s) { return {s.atRegion()}; }
inline SlabRegion inHY(const weighting_function::Slab& s) { return {s.hyRegion()}; }
inline SlabRegion inCG(const weighting_function::Slab& s) { return {s.cgRegion()}; }
}  // namespace pred

// ---- halo ---------------------------------------------------------------------------------
namespace communication
{
/// communication::GhostLayer (GhostLayer.hpp:28-56)
class GhostLayer
{
public:
    GhostLayer() { impl::check(mrmd_ghost_layer_create(context(), &h_), "mrmd_ghost_layer_create"); }
    GhostLayer(const GhostLayer&) = delete;
    ~GhostLayer() { mrmd_ghost_layer_destroy(h_); }
    void exchangeRealAtoms(data::Atoms& atoms, const data::Subdomain& s)
    {
        auto b = s.box();
        impl::check(mrmd_halo_wrap(atoms.sync(), &b), "exchangeRealAtoms");
    }
    void createGhostAtoms(data::Atoms& atoms, const data::Subdomain& s)
    {
        auto b = s.box();
        int64_t n = 0;
        impl::check(mrmd_halo_create(h_, atoms.sync(), &b, &n), "createGhostAtoms");
        atoms.pull();
    }
    void updateGhostAtoms(data::Atoms& atoms, const data::Subdomain& s)
    {
        auto b = s.box();
        impl::check(mrmd_halo_update(h_, atoms.sync(), &b), "updateGhostAtoms");
    }
    void contributeBackGhostToReal(data::Atoms& atoms)
    {
        impl::check(mrmd_halo_fold_forces(h_, atoms.sync()), "contributeBackGhostToReal");
    }
    std::vector<idx_t> getCorrespondingRealAtom(idx_t n) const
    {
        std::vector<idx_t> v(static_cast<size_t>(n));
        impl::check(mrmd_halo_get_corresponding(h_, v.data(), n), "correspondingRealAtom");
        return v;
    }
    mrmd_ghost_layer* handle() const { return h_; }

private:
    mrmd_ghost_layer* h_ = nullptr;
};

/// communication::MultiResGhostLayer (MultiResGhostLayer.hpp:29-61)
class MultiResGhostLayer
{
public:
    MultiResGhostLayer() { impl::check(mrmd_ghost_layer_create(context(), &h_), "mrmd_ghost_layer_create"); }
    MultiResGhostLayer(const MultiResGhostLayer&) = delete;
    ~MultiResGhostLayer() { mrmd_ghost_layer_destroy(h_); }
    void exchangeRealAtoms(data::Molecules& m, data::Atoms& a, const data::Subdomain& s)
    {
        auto b = s.box();
        impl::check(mrmd_multires_wrap(m.sync(), a.sync(), &b), "exchangeRealAtoms");
    }
    void createGhostAtoms(data::Molecules& m, data::Atoms& a, const data::Subdomain& s)
    {
        auto b = s.box();
        impl::check(mrmd_multires_halo_create(h_, m.sync(), a.sync(), &b), "createGhostAtoms");
        m.pull();
        a.pull();
    }
    void updateGhostAtoms(data::Atoms& a, const data::Subdomain& s)
    {
        auto b = s.box();
        impl::check(mrmd_halo_update(h_, a.sync(), &b), "updateGhostAtoms");
    }
    void contributeBackGhostToReal(data::Atoms& a)
    {
        impl::check(mrmd_halo_fold_forces(h_, a.sync()), "contributeBackGhostToReal");
    }

private:
    mrmd_ghost_layer* h_ = nullptr;
};
}  // namespace communication

// ---- actions ------------------------------------------------------------------------------
namespace action
{
/// action::LennardJones (LennardJones.hpp:98-206)
class LennardJones
{
public:
    LennardJones(const real_t rc, const real_t& sigma, const real_t& epsilon, const real_t& cappingDistance = 0_r)
        : LennardJones({cappingDistance}, {rc}, {sigma}, {epsilon}, 1, false)
    {
    }
    LennardJones(const std::vector<real_t>& cap, const std::vector<real_t>& rc, const std::vector<real_t>& sigma,
                 const std::vector<real_t>& epsilon, const idx_t& numTypes, const bool isShifted)
    {
        // the reference keeps numTypes_ = 1 for the pair index (LennardJones.cpp:52)
        impl::check(mrmd_lj_create(context(), cap.data(), rc.data(), sigma.data(), epsilon.data(), numTypes,
                                   isShifted ? 1 : 0, 1, &h_),
                    "LennardJones");
    }
    LennardJones(const LennardJones&) = delete;
    ~LennardJones() { mrmd_lj_destroy(h_); }

    void apply(data::Atoms& atoms, HalfVerletList& vl)
    {
        impl::check(mrmd_lj_apply(h_, atoms.sync(), vl.handle(), nullptr), "LennardJones::apply");
    }
    template <class Pred>
    void apply_if(const data::Atoms& atoms, const HalfVerletList& vl, const Pred& p)
    {
        static_assert(HasDesc2<Pred>::value,
                      "apply_if takes a pair-predicate descriptor (pred::either / pred::both of one-position "
                      "descriptors); device lambdas cannot cross the C ABI");
        const mrmd_pred2 d = p.desc2();
        impl::check(mrmd_lj_apply(h_, atoms.sync(), vl.handle(), &d), "LennardJones::apply_if");
    }
    real_t getEnergy() const
    {
        double e, w;
        impl::check(mrmd_lj_energy_virial(h_, &e, &w), "getEnergy");
        return e;
    }
    real_t getVirial() const
    {
        double e, w;
        impl::check(mrmd_lj_energy_virial(h_, &e, &w), "getVirial");
        return w;
    }
    mrmd_lj* handle() const { return h_; }

private:
    mrmd_lj* h_ = nullptr;
};

/// action::VelocityVerlet (VelocityVerlet.hpp:26-31)
class VelocityVerlet
{
public:
    static real_t preForceIntegrate(data::Atoms& atoms, const real_t dt)
    {
        double d = 0;
        impl::check(mrmd_vv_pre(atoms.sync(), dt, &d), "preForceIntegrate");
        return d;
    }
    static void postForceIntegrate(data::Atoms& atoms, const real_t dt)
    {
        impl::check(mrmd_vv_post(atoms.sync(), dt), "postForceIntegrate");
    }
};

/// action::VelocityVerletLangevinThermostat (VelocityVerletLangevinThermostat.hpp:29-61)
class VelocityVerletLangevinThermostat
{
    real_t zeta_, temperature_;
    uint64_t seed_ = 1234, call_ = 0;

public:
    VelocityVerletLangevinThermostat(const real_t& zeta, const real_t& temperature) { set(zeta, temperature); }
    void set(const real_t& zeta, const real_t& temperature)
    {
        zeta_ = zeta;
        temperature_ = temperature;
    }
    void setSeed(uint64_t seed) { seed_ = seed; }
    real_t preForceIntegrate(data::Atoms& atoms, const real_t dt) { return run(atoms, dt, nullptr); }
    void postForceIntegrate(data::Atoms& atoms, const real_t dt) { VelocityVerlet::postForceIntegrate(atoms, dt); }
    template <class Pred>
    real_t preForceIntegrate_apply_if(data::Atoms& atoms, const real_t dt, const Pred& p)
    {
        const mrmd_pred1 d = toPred1(p);
        return run(atoms, dt, &d);
    }

private:
    real_t run(data::Atoms& atoms, real_t dt, const mrmd_pred1* p)
    {
        double d = 0;
        impl::check(mrmd_langevin_pre(atoms.sync(), dt, zeta_, temperature_, seed_, call_++, p, &d),
                    "Langevin preForceIntegrate");
        return d;
    }
};

/// action::UpdateMolecules (UpdateMolecules.hpp:24-70) with the slab weighting
namespace UpdateMolecules
{
inline void update(const data::Molecules& m, const data::Atoms& a, const weighting_function::Slab& w)
{
    impl::check(mrmd_molecules_update(m.sync(), a.sync(), &w.params()), "UpdateMolecules::update");
}
}  // namespace UpdateMolecules

/// action::ContributeMoleculeForceToAtoms (ContributeMoleculeForceToAtoms.cpp:23-48)
namespace ContributeMoleculeForceToAtoms
{
inline void update(const data::Molecules& m, const data::Atoms& a)
{
    impl::check(mrmd_molecules_contribute_force(m.sync(), a.sync()), "ContributeMoleculeForceToAtoms::update");
}
}  // namespace ContributeMoleculeForceToAtoms

/// action::LJ_IdealGas (LJ_IdealGas.hpp:34-104)
class LJ_IdealGas
{
public:
    LJ_IdealGas(const real_t& cap, const real_t& rc, const real_t& sigma, const real_t& epsilon, const bool doShift)
        : LJ_IdealGas({cap}, {rc}, {sigma}, {epsilon}, 1, doShift)
    {
    }
    LJ_IdealGas(const std::vector<real_t>& cap, const std::vector<real_t>& rc, const std::vector<real_t>& sigma,
                const std::vector<real_t>& epsilon, const idx_t numTypes, const bool doShift)
        : numTypes_(numTypes)
    {
        impl::check(mrmd_ljig_create(context(), cap.data(), rc.data(), sigma.data(), epsilon.data(), numTypes,
                                     doShift ? 1 : 0, &h_),
                    "LJ_IdealGas");
    }
    LJ_IdealGas(const LJ_IdealGas&) = delete;
    ~LJ_IdealGas() { mrmd_ljig_destroy(h_); }
    void setCompensationEnergySamplingInterval(const idx_t& i) { sampling_ = i, push(); }
    void setCompensationEnergyUpdateInterval(const idx_t& i) { update_ = i, push(); }
    real_t run(data::Molecules& m, HalfVerletList& vl, data::Atoms& a)
    {
        double e = 0;
        impl::check(mrmd_ljig_run(h_, m.sync(), vl.handle(), a.sync(), &e), "LJ_IdealGas::run");
        return e;
    }
    data::MultiHistogram getMeanCompensationEnergy() const
    {
        data::MultiHistogram h("meanCompensationEnergy", 0_r, 1_r, 200, numTypes_);
        impl::check(mrmd_ljig_get_histogram(h_, 2, h.data.data()), "getMeanCompensationEnergy");
        return h;
    }

private:
    void push() { impl::check(mrmd_ljig_set_intervals(h_, sampling_, update_), "intervals"); }
    mrmd_lj_idealgas* h_ = nullptr;
    idx_t numTypes_ = 1, sampling_ = 200, update_ = 20000;
};

/// action::ThermodynamicForce (ThermodynamicForce.hpp:32-96)
class ThermodynamicForce
{
public:
    ThermodynamicForce(const std::vector<real_t>& targetDensity, const data::Subdomain& s, const real_t& binWidth,
                       const std::vector<real_t>& modulation, bool enforceSymmetry = false, bool usePeriodicity = false)
        : numTypes_(idx_c(targetDensity.size())), min_(s.minCorner[0]), max_(s.maxCorner[0])
    {
        if (targetDensity.size() != modulation.size()) throw std::invalid_argument("density/modulation sizes differ");
        auto b = s.box();
        impl::check(mrmd_tf_create(context(), targetDensity.data(), modulation.data(), numTypes_, &b, binWidth,
                                   enforceSymmetry, usePeriodicity, &h_),
                    "ThermodynamicForce");
    }
    ThermodynamicForce(real_t targetDensity, const data::Subdomain& s, const real_t& binWidth, real_t modulation,
                       bool enforceSymmetry = false, bool usePeriodicity = false)
        : ThermodynamicForce(std::vector<real_t>{targetDensity}, s, binWidth, std::vector<real_t>{modulation},
                             enforceSymmetry, usePeriodicity)
    {
    }
    ThermodynamicForce(const ThermodynamicForce&) = delete;
    ThermodynamicForce(ThermodynamicForce&& o) noexcept : h_(o.h_), numTypes_(o.numTypes_), min_(o.min_), max_(o.max_)
    {
        o.h_ = nullptr;
    }
    ~ThermodynamicForce() { mrmd_tf_destroy(h_); }
    void sample(data::Atoms& a) { impl::check(mrmd_tf_sample(h_, a.sync()), "sample"); }
    void update(const real_t& sigma, const real_t& range) { impl::check(mrmd_tf_update(h_, sigma, range, nullptr), "update"); }
    template <class Pred>
    void update_if(const real_t& sigma, const real_t& range, const Pred& p)
    {
        const mrmd_pred1 d = toPred1(p);
        impl::check(mrmd_tf_update(h_, sigma, range, &d), "update_if");
    }
    void apply(const data::Atoms& a) const { impl::check(mrmd_tf_apply(h_, a.sync(), nullptr), "apply"); }
    template <class Pred>
    void apply_if(const data::Atoms& a, const Pred& p) const
    {
        const mrmd_pred1 d = toPred1(p);
        impl::check(mrmd_tf_apply(h_, a.sync(), &d), "apply_if");
    }
    template <class Pred>
    void applyInterpolated_if(const data::Atoms& a, const Pred& p) const
    {
        const mrmd_pred1 d = toPred1(p);
        impl::check(mrmd_tf_apply_interpolated(h_, a.sync(), &d), "applyInterpolated_if");
    }
    data::MultiHistogram getForce() const { return get(0, "thermodynamic-force"); }
    /// the force of one atom type over the bins (ThermodynamicForce.hpp:49-52)
    std::vector<real_t> getForce(const idx_t& typeId) const
    {
        const auto f = getForce();
        std::vector<real_t> out(static_cast<size_t>(f.numBins));
        for (idx_t b = 0; b < f.numBins; ++b) out[size_t(b)] = f(b, typeId);
        return out;
    }
    data::MultiHistogram getDensityProfile() const { return get(1, "density-profile"); }
    void setForce(const data::MultiHistogram& f) const { impl::check(mrmd_tf_set_force(h_, f.data.data()), "setForce"); }
    idx_t getNumberOfDensityProfileSamples() const
    {
        int64_t s = 0;
        impl::check(mrmd_tf_samples(h_, &s), "samples");
        return s;
    }
    std::vector<real_t> getMuLeft() const { return mu(0); }
    std::vector<real_t> getMuRight() const { return mu(1); }

private:
    data::MultiHistogram get(int which, const char* label) const
    {
        int64_t nb = 0;
        impl::check(mrmd_tf_num_bins(h_, &nb), "bins");
        data::MultiHistogram h(label, min_, max_, nb, numTypes_);
        impl::check(mrmd_tf_get(h_, which, h.data.data()), "get");
        return h;
    }
    std::vector<real_t> mu(int side) const
    {
        std::vector<real_t> l(static_cast<size_t>(numTypes_)), r(static_cast<size_t>(numTypes_));
        impl::check(mrmd_tf_mu(h_, l.data(), r.data()), "mu");
        return side == 0 ? l : r;
    }
    mrmd_thermo_force* h_ = nullptr;
    idx_t numTypes_;
    real_t min_, max_;
};
}  // namespace action

namespace data
{
/// createGrid (MultiHistogram.cpp:214-227): the bin centres
inline std::vector<real_t> createGrid(const MultiHistogram& h)
{
    std::vector<real_t> g(static_cast<size_t>(h.numBins));
    for (idx_t b = 0; b < h.numBins; ++b) g[size_t(b)] = h.getBinPosition(b);
    return g;
}
}  // namespace data

namespace action
{
/// action::BerendsenThermostat (BerendsenThermostat.hpp:22-36)
namespace BerendsenThermostat
{
inline void apply(data::Atoms& atoms, const real_t& currentTemperature, const real_t& targetTemperature,
                  const real_t& gamma)
{
    impl::check(mrmd_berendsen_apply(atoms.sync(), currentTemperature, targetTemperature, gamma),
                "BerendsenThermostat::apply");
}
}  // namespace BerendsenThermostat
}  // namespace action

// ---- analysis / io ----------------------------------------------------------------------
namespace analysis
{
/// analysis::getKineticEnergy (KineticEnergy.hpp:26-39)
inline real_t getKineticEnergy(data::Atoms& atoms)
{
    double e = 0;
    impl::check(mrmd_kinetic_energy(atoms.sync(), &e), "getKineticEnergy");
    return e;
}
inline real_t getMeanKineticEnergy(data::Atoms& atoms)
{
    return getKineticEnergy(atoms) / real_c(atoms.numLocalAtoms);
}
}  // namespace analysis

namespace io
{
/// io::restoreAtoms (mrmd/io/RestoreTXT.cpp:24-74): "x y z" per line, mass 1, type 0
inline data::Atoms restoreAtoms(const std::string& filename)
{
    std::ifstream fin(filename);
    if (!fin.is_open())
    {
        std::cerr << "Could not open file: " << filename << std::endl;
        exit(EXIT_FAILURE);
    }
    data::HostAtoms h(0);
    idx_t n = 0;
    while (!fin.eof())
    {
        double x, y, z;
        fin >> x >> y >> z;
        if (fin.eof()) break;
        if (std::isnan(x) || std::isnan(y) || std::isnan(z))
        {
            std::cout << "invalid position: " << x << " " << y << " " << z << std::endl;
            exit(EXIT_FAILURE);
        }
        h.resize(size_t(n + 1));
        h.getPos()(n, 0) = x;
        h.getPos()(n, 1) = y;
        h.getPos()(n, 2) = z;
        h.getMass()(n) = 1_r;
        h.getType()(n) = 0;
        ++n;
    }
    h.numLocalAtoms = n;
    data::Atoms a(n);
    data::deep_copy(a, h);
    return a;
}

/// io::restoreGRO (mrmd/io/RestoreGRO.cpp:24-147): a comment line, the atom count, one
/// fixed-width line per atom (residue number, residue name, atom name, atom number,
/// position and optionally velocity, scanned as "%5d%5s%5s%5d%8lf%8lf%8lf%8lf%8lf%8lf"),
/// then the box edges, which set the subdomain's diameter from its minimum corner (ghost
/// layer thickness kept).  Mass and relative mass 1, type 0, force 0; a line without
/// velocities gives velocity 0 (with one warning).  Files with ghost atoms are refused.
inline void restoreGRO(const std::string& filename,
                       data::Subdomain& subdomain,
                       data::Atoms& atoms,
                       bool containsGhostAtoms = false)
{
    std::ifstream fin(filename);
    if (!fin.is_open())
    {
        std::cerr << "Could not open file: " << filename << std::endl;
        exit(EXIT_FAILURE);
    }
    if (containsGhostAtoms)
    {
        std::cout << "Reading a file with ghost atoms is not supported." << std::endl;
        exit(EXIT_FAILURE);
    }
    std::string line;
    std::getline(fin, line);  // comment
    std::getline(fin, line);
    int count = 0;
    if (std::sscanf(line.c_str(), "%d", &count) != 1 || count <= 0)
    {
        std::cerr << "Invalid atom count in file " << filename << std::endl;
        exit(EXIT_FAILURE);
    }
    data::HostAtoms h{idx_t(count)};
    bool noVelocities = false;
    idx_t n = 0;
    while (n < count && std::getline(fin, line))
    {
        int resNr = 0, atomNr = 0;
        char resName[6] = {}, atomName[6] = {};
        double v[6] = {0, 0, 0, 0, 0, 0};
        const int fields = std::sscanf(line.c_str(), "%5d%5s%5s%5d%8lf%8lf%8lf%8lf%8lf%8lf", &resNr, resName,
                                       atomName, &atomNr, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5]);
        if (fields != 10)
        {
            if (fields != 7)
            {
                std::cerr << "Invalid GRO atom line in file " << filename
                          << ": expected 7 fields for coordinates (or 10 fields including velocities), but parsed "
                          << fields << " fields from line: \"" << line << "\"" << std::endl;
                exit(EXIT_FAILURE);
            }
            noVelocities = true;
            v[3] = v[4] = v[5] = 0.0;
        }
        for (int d = 0; d < 3; ++d)
        {
            h.getPos()(n, d) = v[d];
            h.getVel()(n, d) = v[3 + d];
        }
        h.getMass()(n) = 1_r;
        h.getRelativeMass()(n) = 1_r;
        h.getType()(n) = 0;
        ++n;
    }
    Vector3D diam{};
    fin >> diam[0] >> diam[1] >> diam[2];
    if (noVelocities)
        std::cout << "Warning: Some lines in file " << filename
                  << " do not contain velocities. Respective velocities have been set to zero, but this may lead to "
                     "unexpected behavior."
                  << std::endl;
    subdomain = data::Subdomain(subdomain.minCorner,
                                {subdomain.minCorner[0] + diam[0], subdomain.minCorner[1] + diam[1],
                                 subdomain.minCorner[2] + diam[2]},
                                subdomain.ghostLayerThickness);
    h.numLocalAtoms = n;
    h.numGhostAtoms = 0;
    atoms = data::Atoms(n);
    data::deep_copy(atoms, h);
    atoms.numLocalAtoms = n;
    atoms.numGhostAtoms = 0;
}
}  // namespace io
}  // namespace mrmd
