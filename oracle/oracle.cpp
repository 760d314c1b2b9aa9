// TEST INFRASTRUCTURE ONLY - NOT PART OF THE PRODUCT.
//
// CPU restatement of the MRMD hot path.  Every function cites the reference file:line it
// follows (paths relative to the XzzX/mrmd source tree).  It is the parity checker of the
// HIP library and the timed host-core baseline; see oracle.hpp for the layout conventions.
//
// Parity status of each part (what pins it):
//  * LJ potential/force/energy: LennardJones.test.cpp closed form and the ESPResSo++ energy
//    -94795.927 of tests/LennardJones/LennardJones.cpp; F = -grad E by finite differences.
//  * cell binning / half Verlet list: ESPResSo++ pair count 1,310,403 and brute force.
//  * ghosts / wrap / ghost update / fold: GhostExchange.test.cpp (18 / 98 ghosts, 108 pairs),
//    PeriodicMapping.test.cpp, UpdateGhostAtoms.test.cpp, AccumulateForce.test.cpp.
//  * velocity Verlet: VelocityVerlet.test.cpp; Langevin with a false predicate:
//    VelocityVerletLangevinThermostat.test.cpp; Langevin noise: Philox4x32-10 known-answer
//    vectors (Random123) plus moment tests - the reference's Kokkos XorShift1024 stream is
//    parity unpinned (statistical agreement only).
//  * AdResS LJ/ideal gas: LJ_IdealGas.test.cpp 4-atom forces (grad lambda = 0 there); the
//    drift term by the gradient identity (atom forces after ContributeMoleculeForceToAtoms
//    and the ghost fold = -grad of the returned weighted energy, central differences, nu = 2
//    so lambda_mod != lambda, diatomic molecules and a periodic slab across the boundary);
//    compensation samples, counts, the running mean and the mean compensation force by a
//    literal fixture (equilateral triangle at the LJ minimum: every V_ij = -epsilon/2) and a
//    closed-form five-molecule case over all three regions and two types
//    (tests/test_oracle_adress_pins.py).  The reference's compensation counter is a data
//    race (LJ_IdealGas.cpp:204-208); this restatement counts atomically.
//  * thermodynamic force / histograms: ThermodynamicForce.test.cpp, MultiHistogram.test.cpp.
//  * slab weighting / regions / molecule update / force distribution: Slab.test.cpp,
//    CheckRegion.test.cpp, UpdateMolecules.test.cpp, ContributeMoleculeForceToAtoms.test.cpp,
//    plus finite differences of grad lambda.
//  * multi-resolution ghosts: MultiResPeriodicGhostExchange.test.cpp counts,
//    MultiResRealAtomsExchange.test.cpp.
#include "oracle.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

namespace
{
constexpr double PI = 3.14159265358979323846;

// ---------------------------------------------------------------------------------------
// Subdomain (mrmd/data/Subdomain.hpp:41-62)
struct Sub
{
    double min[3], max[3], g[3], minGhost[3], maxGhost[3], minInner[3], maxInner[3], diam[3];
};

Sub makeSub(const or_box& b)
{
    Sub s;
    for (int d = 0; d < 3; ++d)
    {
        s.min[d] = b.min[d];
        s.max[d] = b.max[d];
        s.g[d] = b.ghost[d];
        s.minGhost[d] = s.min[d] - s.g[d];
        s.maxGhost[d] = s.max[d] + s.g[d];
        s.minInner[d] = s.min[d] + s.g[d];
        s.maxInner[d] = s.max[d] - s.g[d];
        s.diam[d] = s.max[d] - s.min[d];
    }
    return s;
}

// ---------------------------------------------------------------------------------------
// containers (mrmd/data/Atoms.hpp:37-54, mrmd/data/Molecules.hpp:31-48)
struct Atom
{
    double pos[3];
    double vel[3];
    double force[3];
    int64_t type;
    double mass;
    double charge;
    double relMass;
};
static_assert(sizeof(Atom) == 104, "reference AoS record is 104 bytes");

struct Molecule
{
    double pos[3];
    double force[3];
    double lambda;
    double modLambda;
    double grad[3];
    int64_t offset;
    int64_t numAtoms;
};
}  // namespace

struct OAtoms
{
    std::vector<Atom> a;
    int64_t nlocal = 0;
    int64_t nghost = 0;
};

struct OMolecules
{
    std::vector<Molecule> m;
    int64_t nlocal = 0;
    int64_t nghost = 0;
};

namespace
{
// ---------------------------------------------------------------------------------------
// predicates (descriptor form of the reference's predicate functors:
// util/IsInSymmetricSlab.hpp:34-50, weighting_function/Slab.hpp:44-71)
bool evalPred1(const or_pred1& p, double x, double y, double z)
{
    const double c[3] = {x, y, z};
    const double v = c[p.axis];
    switch (p.kind)
    {
        case 0:
            return true;
        case 1:
            return false;
        case 2:
            return v > p.a && v < p.b;
        case 3:
        {
            auto absDx = std::abs(v - p.a);
            return absDx >= p.b - p.tol && absDx <= p.c + p.tol;
        }
        case 4:
            return std::abs(v - p.a) < p.b;
        case 5:
        {
            auto absDx = std::abs(v - p.a);
            if (absDx < p.b) return false;
            return absDx < p.b + p.c;
        }
        case 6:
            return !(std::abs(v - p.a) < p.b + p.c);
        case 7:
            return v > p.a;
        case 8:
            return v < p.a;
        default:
            return false;
    }
}

bool evalPred2(const or_pred2& p, const double* a, const double* b)
{
    switch (p.op)
    {
        case 0:
            return true;
        case 1:
            return evalPred1(p.first, a[0], a[1], a[2]) && evalPred1(p.second, b[0], b[1], b[2]);
        case 2:
            return evalPred1(p.first, a[0], a[1], a[2]) || evalPred1(p.second, b[0], b[1], b[2]);
        default:
            return false;
    }
}

template <class T>
void atomicAdd(T& target, T val)
{
#pragma omp atomic
    target += val;
}

// ---------------------------------------------------------------------------------------
// Cabana Impl::CartesianGrid semantics (third-party, Cabana 0.7, not vendored; pinned by
// the ESPResSo++ pair count): n = floor((max-min)*(1/delta)), dx = (max-min)/n, cell index
// floor((x-min)*(1/dx)) with the outer boundary folded into the last cell,
// cardinal index (i*ny + j)*nz + k.
struct Grid
{
    double min[3], max[3], dx[3], rdx[3];
    int n[3];

    void init(const double* mn, const double* mx, const double* delta)
    {
        for (int d = 0; d < 3; ++d)
        {
            min[d] = mn[d];
            max[d] = mx[d];
            n[d] = static_cast<int>(std::floor((mx[d] - mn[d]) * (1.0 / delta[d])));
            if (n[d] < 1) n[d] = 1;
            dx[d] = (mx[d] - mn[d]) / n[d];
            rdx[d] = 1.0 / dx[d];
        }
    }
    int locate1(double x, int d) const
    {
        int c = static_cast<int>(std::floor((x - min[d]) * rdx[d]));
        if (c == n[d]) c = n[d] - 1;
        // outside the grid is undefined behaviour in the reference; clamp instead
        if (c < 0) c = 0;
        if (c >= n[d]) c = n[d] - 1;
        return c;
    }
    int cardinal(int i, int j, int k) const { return (i * n[1] + j) * n[2] + k; }
    int total() const { return n[0] * n[1] * n[2]; }
    double minDistSqr(const double* p, int i, int j, int k) const
    {
        const int ijk[3] = {i, j, k};
        double r2 = 0.0;
        for (int d = 0; d < 3; ++d)
        {
            double c = min[d] + (ijk[d] + 0.5) * dx[d];
            double r = std::abs(p[d] - c) - 0.5 * dx[d];
            r = (r > 0.0) ? r : 0.0;
            r2 = r2 + r * r;
        }
        return r2;
    }
};

// LJ precomputed values (mrmd/action/LennardJones.hpp:27-39)
struct Pre
{
    double ff1, ff2, ef1, ef2, rcSqr, capDist, capDistSqr, capCoeff, shift, eCap;
};

struct FE
{
    double ff;
    double e;
};

// mrmd/action/LennardJones.hpp:52-78
inline FE computeForceAndEnergy(const Pre& p, double distSqr)
{
    FE r;
    if (distSqr >= p.capDistSqr)
    {
        double frac2 = 1.0 / distSqr;
        double frac6 = frac2 * frac2 * frac2;
        r.ff = frac6 * (p.ff1 * frac6 - p.ff2) * frac2;
        r.e = frac6 * (p.ef1 * frac6 - p.ef2) - p.shift;
        return r;
    }
    double dist = std::sqrt(distSqr);
    r.ff = p.capCoeff / dist;
    r.e = p.eCap - (dist - p.capDist) * p.capCoeff - p.shift;
    return r;
}

// ---------------------------------------------------------------------------------------
// counter-based noise: Philox4x32-10 (Salmon et al., SC'11) + Marsaglia polar normals.
// This replaces the reference's Kokkos Random_XorShift1024_Pool(1234)
// (VelocityVerletLangevinThermostat.hpp:32); the reference stream itself cannot be
// reproduced, so Langevin parity against the reference is statistical only.
void philox4x32_10(const uint32_t ctrIn[4], const uint32_t keyIn[2], uint32_t out[4])
{
    uint32_t c[4] = {ctrIn[0], ctrIn[1], ctrIn[2], ctrIn[3]};
    uint32_t k0 = keyIn[0], k1 = keyIn[1];
    for (int round = 0; round < 10; ++round)
    {
        uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
        uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
        uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    for (int i = 0; i < 4; ++i) out[i] = c[i];
}

// uniform in (-1, 1) from 52 random bits; every step is exact in binary64
inline double symUniform(uint32_t a, uint32_t b)
{
    uint64_t m = (uint64_t(a) << 20) ^ uint64_t(b >> 12);
    return (double(m) + 0.5) * 0x1p-51 - 1.0;
}

// natural logarithm for x in (0, 1] using only +,-,*,/ (bit-reproducible on any IEEE
// binary64 machine): x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(t),
// t = (m-1)/(m+1), series to t^25.
inline double detLog(double x)
{
    uint64_t bits;
    std::memcpy(&bits, &x, 8);
    int e = int((bits >> 52) & 0x7ff) - 1023;
    uint64_t mb = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    std::memcpy(&m, &mb, 8);
    if (m > 1.4142135623730951)
    {
        m = m * 0.5;
        e += 1;
    }
    double t = (m - 1.0) / (m + 1.0);
    double t2 = t * t;
    double p = 1.0 / 25.0;
    p = p * t2 + 1.0 / 23.0;
    p = p * t2 + 1.0 / 21.0;
    p = p * t2 + 1.0 / 19.0;
    p = p * t2 + 1.0 / 17.0;
    p = p * t2 + 1.0 / 15.0;
    p = p * t2 + 1.0 / 13.0;
    p = p * t2 + 1.0 / 11.0;
    p = p * t2 + 1.0 / 9.0;
    p = p * t2 + 1.0 / 7.0;
    p = p * t2 + 1.0 / 5.0;
    p = p * t2 + 1.0 / 3.0;
    p = p * t2 + 1.0;
    return double(e) * 0.6931471805599453 + 2.0 * t * p;
}

// exp(x) for x <= 0 using only +,-,*,/ : x = k ln2 + r (Cody-Waite split of ln2),
// Horner form of the Taylor series of e^r to degree 13, then scaling by 2^k.
inline double detExp(double x)
{
    if (x < -700.0) return 0.0;
    if (x == 0.0) return 1.0;
    double kd = std::floor(x * 1.4426950408889634 + 0.5);
    double r = (x - kd * 6.93147180369123816490e-01) - kd * 1.90821492927058770002e-10;
    double p = 1.0;
    for (int n = 13; n >= 1; --n) p = 1.0 + (p * r) / double(n);
    int k = int(kd);
    uint64_t sb = uint64_t(k + 1023) << 52;
    double scale;
    std::memcpy(&scale, &sb, 8);
    return p * scale;
}

// three standard normals for (key, id, call)
void normals3(uint64_t seed, uint64_t id, uint64_t call, double out[3])
{
    const uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    int have = 0;
    for (uint32_t round = 0; have < 3; ++round)
    {
        const uint32_t ctr[4] = {uint32_t(id), uint32_t(call), uint32_t(call >> 32), round};
        uint32_t r[4];
        philox4x32_10(ctr, key, r);
        double u = symUniform(r[0], r[1]);
        double v = symUniform(r[2], r[3]);
        double s = u * u + v * v;
        if (!(s < 1.0 && s > 0.0)) continue;
        double f = std::sqrt(-2.0 * detLog(s) / s);
        out[have++] = u * f;
        if (have < 3) out[have++] = v * f;
    }
}

// ---------------------------------------------------------------------------------------
// MultiHistogram (mrmd/data/MultiHistogram.hpp:29-92, MultiHistogram.cpp:27-227)
struct Hist
{
    double min = 0, max = 1;
    int64_t numBins = 0, numHist = 0;
    double binSize = 1, invBinSize = 1;
    std::vector<double> data;  // [bin * numHist + hist]

    void init(double mn, double mx, int64_t nb, int64_t nh)
    {
        min = mn;
        max = mx;
        numBins = nb;
        numHist = nh;
        binSize = (max - min) / double(numBins);
        invBinSize = 1.0 / binSize;
        data.assign(size_t(nb * nh), 0.0);
    }
    double& at(int64_t b, int64_t h) { return data[size_t(b * numHist + h)]; }
    double at(int64_t b, int64_t h) const { return data[size_t(b * numHist + h)]; }
    int64_t getBin(double v) const
    {
        auto bin = int64_t(std::floor((v - min) * invBinSize));
        if (bin < 0) bin = -1;
        if (bin >= numBins) bin = -1;
        return bin;
    }
    double getBinPosition(int64_t b) const { return min + (double(b) + 0.5) * binSize; }
    void scale(double f)
    {
        for (auto& v : data) v *= f;
    }
    void makeSymmetric()
    {
        auto maxIdx = numBins - 1;
        for (int64_t i = 0; i < numBins / 2; ++i)
            for (int64_t h = 0; h < numHist; ++h)
            {
                auto val = 0.5 * (at(i, h) + at(maxIdx - i, h));
                at(i, h) = val;
                at(maxIdx - i, h) = val;
            }
    }
};

// MultiHistogram.cpp:113-161
Hist gradient(const Hist& in, bool periodic)
{
    const auto invSp = in.invBinSize;
    const auto invDSp = 0.5 * in.invBinSize;
    Hist g;
    g.init(in.min, in.max, in.numBins, in.numHist);
    for (int64_t i = 0; i < in.numBins; ++i)
        for (int64_t h = 0; h < in.numHist; ++h)
        {
            if (i == 0)
            {
                g.at(i, h) = periodic ? (in.at(i + 1, h) - in.at(in.numBins - 1, h)) * invDSp
                                      : (in.at(i + 1, h) - in.at(i, h)) * invSp;
                continue;
            }
            if (i == in.numBins - 1)
            {
                g.at(i, h) = periodic ? (in.at(0, h) - in.at(i - 1, h)) * invDSp
                                      : (in.at(i, h) - in.at(i - 1, h)) * invSp;
                continue;
            }
            g.at(i, h) = (in.at(i + 1, h) - in.at(i - 1, h)) * invDSp;
        }
    return g;
}

// MultiHistogram.cpp:163-212
Hist smoothen(const Hist& in, double sigma, double range, bool periodic)
{
    const auto invSigma = 1.0 / sigma;
    const int64_t delta = int64_t(int(range * sigma * in.invBinSize));
    Hist out;
    out.init(in.min, in.max, in.numBins, in.numHist);
    for (int64_t b = 0; b < in.numBins; ++b)
        for (int64_t h = 0; h < in.numHist; ++h)
        {
            double norm = 0.0;
            int64_t jmin = b - delta, jmax = b + delta;
            if (!periodic)
            {
                jmin = std::max(int64_t(0), jmin);
                jmax = std::min(in.numBins - 1, jmax);
            }
            for (auto j = jmin; j <= jmax; ++j)
            {
                auto mj = j;
                if (periodic)
                {
                    if (mj < 0) mj += in.numBins;
                    if (mj >= in.numBins) mj -= in.numBins;
                }
                double t = double(b - j) * in.binSize * invSigma;
                const double eFunc = std::exp(-(t * t));
                norm += eFunc;
                out.at(b, h) += in.at(mj, h) * eFunc;
            }
            out.at(b, h) /= norm;
        }
    return out;
}
}  // namespace

// =========================================================================================
// C ABI for ctypes
extern "C" {

void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int or_get_threads() { return omp_get_max_threads(); }

// ---- deterministic math / RNG (new code; pinned by known-answer vectors and moments) ----
void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out)
{
    philox4x32_10(ctr, key, out);
}
double or_det_log(double x) { return detLog(x); }
double or_det_exp(double x) { return detExp(x); }
void or_normals3(uint64_t seed, uint64_t id, uint64_t call, double* out)
{
    normals3(seed, id, call, out);
}

// ---- atoms container ----------------------------------------------------------------
OAtoms* or_atoms_new() { return new OAtoms(); }
void or_atoms_free(OAtoms* a) { delete a; }
void or_atoms_resize(OAtoms* a, int64_t n) { a->a.resize(size_t(n), Atom{}); }
int64_t or_atoms_size(const OAtoms* a) { return int64_t(a->a.size()); }
void or_atoms_set_counts(OAtoms* a, int64_t nl, int64_t ng)
{
    a->nlocal = nl;
    a->nghost = ng;
}
int64_t or_atoms_nlocal(const OAtoms* a) { return a->nlocal; }
int64_t or_atoms_nghost(const OAtoms* a) { return a->nghost; }

// field: 0 pos, 1 vel, 2 force (n x 3 doubles), 3 type (int64), 4 mass, 5 charge, 6 relMass
void or_atoms_set(OAtoms* a, int field, const void* src, int64_t n)
{
    if (int64_t(a->a.size()) < n) a->a.resize(size_t(n), Atom{});
    const double* d = static_cast<const double*>(src);
    const int64_t* l = static_cast<const int64_t*>(src);
    for (int64_t i = 0; i < n; ++i)
    {
        auto& at = a->a[size_t(i)];
        switch (field)
        {
            case 0:
                for (int k = 0; k < 3; ++k) at.pos[k] = d[3 * i + k];
                break;
            case 1:
                for (int k = 0; k < 3; ++k) at.vel[k] = d[3 * i + k];
                break;
            case 2:
                for (int k = 0; k < 3; ++k) at.force[k] = d[3 * i + k];
                break;
            case 3:
                at.type = l[i];
                break;
            case 4:
                at.mass = d[i];
                break;
            case 5:
                at.charge = d[i];
                break;
            case 6:
                at.relMass = d[i];
                break;
        }
    }
}

void or_atoms_get(const OAtoms* a, int field, void* dst, int64_t n)
{
    double* d = static_cast<double*>(dst);
    int64_t* l = static_cast<int64_t*>(dst);
    for (int64_t i = 0; i < n && i < int64_t(a->a.size()); ++i)
    {
        const auto& at = a->a[size_t(i)];
        switch (field)
        {
            case 0:
                for (int k = 0; k < 3; ++k) d[3 * i + k] = at.pos[k];
                break;
            case 1:
                for (int k = 0; k < 3; ++k) d[3 * i + k] = at.vel[k];
                break;
            case 2:
                for (int k = 0; k < 3; ++k) d[3 * i + k] = at.force[k];
                break;
            case 3:
                l[i] = at.type;
                break;
            case 4:
                d[i] = at.mass;
                break;
            case 5:
                d[i] = at.charge;
                break;
            case 6:
                d[i] = at.relMass;
                break;
        }
    }
}

// Atoms::setForce (Atoms.hpp:72) over all records
void or_atoms_set_force(OAtoms* a, double v)
{
    for (auto& at : a->a)
        for (int k = 0; k < 3; ++k) at.force[k] = v;
}

// ---- molecules container ---------------------------------------------------------------
OMolecules* or_mol_new() { return new OMolecules(); }
void or_mol_free(OMolecules* m) { delete m; }
void or_mol_resize(OMolecules* m, int64_t n) { m->m.resize(size_t(n), Molecule{}); }
int64_t or_mol_size(const OMolecules* m) { return int64_t(m->m.size()); }
void or_mol_set_counts(OMolecules* m, int64_t nl, int64_t ng)
{
    m->nlocal = nl;
    m->nghost = ng;
}
int64_t or_mol_nlocal(const OMolecules* m) { return m->nlocal; }
int64_t or_mol_nghost(const OMolecules* m) { return m->nghost; }

// field: 0 pos, 1 force, 4 gradLambda (n x 3); 2 lambda, 3 modLambda; 5 offset, 6 numAtoms
void or_mol_set(OMolecules* m, int field, const void* src, int64_t n)
{
    if (int64_t(m->m.size()) < n) m->m.resize(size_t(n), Molecule{});
    const double* d = static_cast<const double*>(src);
    const int64_t* l = static_cast<const int64_t*>(src);
    for (int64_t i = 0; i < n; ++i)
    {
        auto& mo = m->m[size_t(i)];
        switch (field)
        {
            case 0:
                for (int k = 0; k < 3; ++k) mo.pos[k] = d[3 * i + k];
                break;
            case 1:
                for (int k = 0; k < 3; ++k) mo.force[k] = d[3 * i + k];
                break;
            case 2:
                mo.lambda = d[i];
                break;
            case 3:
                mo.modLambda = d[i];
                break;
            case 4:
                for (int k = 0; k < 3; ++k) mo.grad[k] = d[3 * i + k];
                break;
            case 5:
                mo.offset = l[i];
                break;
            case 6:
                mo.numAtoms = l[i];
                break;
        }
    }
}

void or_mol_get(const OMolecules* m, int field, void* dst, int64_t n)
{
    double* d = static_cast<double*>(dst);
    int64_t* l = static_cast<int64_t*>(dst);
    for (int64_t i = 0; i < n && i < int64_t(m->m.size()); ++i)
    {
        const auto& mo = m->m[size_t(i)];
        switch (field)
        {
            case 0:
                for (int k = 0; k < 3; ++k) d[3 * i + k] = mo.pos[k];
                break;
            case 1:
                for (int k = 0; k < 3; ++k) d[3 * i + k] = mo.force[k];
                break;
            case 2:
                d[i] = mo.lambda;
                break;
            case 3:
                d[i] = mo.modLambda;
                break;
            case 4:
                for (int k = 0; k < 3; ++k) d[3 * i + k] = mo.grad[k];
                break;
            case 5:
                l[i] = mo.offset;
                break;
            case 6:
                l[i] = mo.numAtoms;
                break;
        }
    }
}

void or_mol_set_force(OMolecules* m, double v)
{
    for (auto& mo : m->m)
        for (int k = 0; k < 3; ++k) mo.force[k] = v;
}

// MoleculesFromAtoms.cpp:19-39
OMolecules* or_create_molecule_for_each_atom(const OAtoms* a)
{
    auto* m = new OMolecules();
    auto size = a->nlocal + a->nghost;
    m->m.assign(size_t(2 * size), Molecule{});
    for (int64_t i = 0; i < size; ++i)
    {
        m->m[size_t(i)].offset = i;
        m->m[size_t(i)].numAtoms = 1;
    }
    m->nlocal = a->nlocal;
    m->nghost = a->nghost;
    return m;
}

// ---- periodic halo ---------------------------------------------------------------------
struct OGhost
{
    std::vector<int64_t> corr;  ///< correspondingRealAtom (GhostExchange.hpp:39-40)
};
OGhost* or_ghost_new() { return new OGhost(); }
void or_ghost_free(OGhost* g) { delete g; }
void or_ghost_get_corr(const OGhost* g, int64_t* out, int64_t n)
{
    for (int64_t i = 0; i < n && i < int64_t(g->corr.size()); ++i) out[i] = g->corr[size_t(i)];
}
int64_t or_ghost_corr_size(const OGhost* g) { return int64_t(g->corr.size()); }
void or_ghost_set_corr(OGhost* g, const int64_t* in, int64_t n)
{
    g->corr.assign(in, in + n);
}

// PeriodicMapping.cpp:30-58
void or_map_into_domain(OAtoms* a, const or_box* box)
{
    auto s = makeSub(*box);
#pragma omp parallel for
    for (int64_t i = 0; i < a->nlocal; ++i)
        for (int d = 0; d < 3; ++d)
        {
            auto& x = a->a[size_t(i)].pos[d];
            if (s.max[d] <= x)
            {
                x -= s.diam[d];
                x = std::max(x, s.min[d]);
            }
            if (x < s.min[d])
            {
                x += s.diam[d];
                if (s.max[d] <= x) x = s.min[d];
            }
        }
}

// GhostExchange.cpp:183-187
void or_ghost_reset(OGhost* g, const OAtoms* a) { g->corr.assign(size_t(a->nlocal), -1); }

// GhostExchange.cpp:59-169 (ordered selection = the reference's parallel_scan)
void or_ghost_create_axis(OGhost* g, OAtoms* a, const or_box* box, int axis)
{
    auto s = makeSub(*box);
    const int64_t total = a->nlocal + a->nghost;
    std::vector<int64_t> P, N;
    for (int64_t i = 0; i < total; ++i)
    {
        const double x = a->a[size_t(i)].pos[axis];
        if (x < s.minInner[axis]) P.push_back(i);
        if (x >= s.maxInner[axis]) N.push_back(i);
    }
    const int64_t nP = int64_t(P.size()), nN = int64_t(N.size());
    a->a.resize(size_t(total + nP + nN));
    if (int64_t(g->corr.size()) < total + nP + nN) g->corr.resize(size_t(total + nP + nN), 0);
    auto resolve = [&](int64_t src)
    {
        auto real = src;
        while (g->corr[size_t(real)] != -1) real = g->corr[size_t(real)];
        return real;
    };
    for (int64_t k = 0; k < nP; ++k)
    {
        auto gi = total + k;
        a->a[size_t(gi)] = a->a[size_t(P[size_t(k)])];
        a->a[size_t(gi)].pos[axis] += s.diam[axis];
        g->corr[size_t(gi)] = resolve(P[size_t(k)]);
    }
    for (int64_t k = 0; k < nN; ++k)
    {
        auto gi = total + nP + k;
        a->a[size_t(gi)] = a->a[size_t(N[size_t(k)])];
        a->a[size_t(gi)].pos[axis] -= s.diam[axis];
        g->corr[size_t(gi)] = resolve(N[size_t(k)]);
    }
    a->nghost += nP + nN;
}

// GhostExchange.cpp:171-181
void or_ghost_create_xyz(OGhost* g, OAtoms* a, const or_box* box)
{
    or_ghost_reset(g, a);
    a->nghost = 0;
    for (int d = 0; d < 3; ++d) or_ghost_create_axis(g, a, box, d);
}

// UpdateGhostAtoms.cpp:31-68
void or_ghost_update(const OGhost* g, OAtoms* a, const or_box* box)
{
    auto s = makeSub(*box);
#pragma omp parallel for
    for (int64_t i = a->nlocal; i < a->nlocal + a->nghost; ++i)
    {
        auto real = g->corr[size_t(i)];
        auto& p = a->a[size_t(i)].pos;
        const auto& q = a->a[size_t(real)].pos;
        double dx[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
        p[0] = q[0];
        p[1] = q[1];
        p[2] = q[2];
        double delta[3] = {0.1 * s.diam[0], 0.1 * s.diam[1], 0.1 * s.diam[2]};
        for (int d = 0; d < 3; ++d)
            if (dx[d] > +delta[d]) p[d] += s.diam[d];
        for (int d = 0; d < 3; ++d)
            if (dx[d] < -delta[d]) p[d] -= s.diam[d];
    }
}

// AccumulateForce.cpp:25-47
void or_ghost_fold(const OGhost* g, OAtoms* a)
{
#pragma omp parallel for
    for (int64_t i = a->nlocal; i < a->nlocal + a->nghost; ++i)
    {
        auto real = g->corr[size_t(i)];
        if (real == -1) continue;
        for (int d = 0; d < 3; ++d)
        {
            atomicAdd(a->a[size_t(real)].force[d], a->a[size_t(i)].force[d]);
            a->a[size_t(i)].force[d] = 0.0;
        }
    }
}

// ---- linked cell list + permute (Cabana LinkedCellList / Cabana::permute semantics;
//      in-cell order is ascending particle index, i.e. the stable version of Cabana's
//      atomic fill) -------------------------------------------------------------------------
struct OCells
{
    Grid grid;
    int64_t begin = 0, end = 0;
    std::vector<int> cellOf;       ///< cell of particle begin+k
    std::vector<int> counts;       ///< per cell
    std::vector<int64_t> offsets;  ///< per cell (exclusive scan)
    std::vector<int64_t> perm;     ///< sorted slot k -> particle index
};

static void buildCells(OCells& c, const std::vector<Atom>& a, int64_t begin, int64_t end)
{
    c.begin = begin;
    c.end = end;
    const int nc = c.grid.total();
    c.cellOf.assign(size_t(end - begin), 0);
    c.counts.assign(size_t(nc), 0);
    for (int64_t p = begin; p < end; ++p)
    {
        const auto& x = a[size_t(p)].pos;
        int cell = c.grid.cardinal(c.grid.locate1(x[0], 0), c.grid.locate1(x[1], 1),
                                   c.grid.locate1(x[2], 2));
        c.cellOf[size_t(p - begin)] = cell;
        c.counts[size_t(cell)]++;
    }
    c.offsets.assign(size_t(nc) + 1, 0);
    for (int i = 0; i < nc; ++i) c.offsets[size_t(i) + 1] = c.offsets[size_t(i)] + c.counts[size_t(i)];
    std::vector<int64_t> fill(c.offsets.begin(), c.offsets.end() - 1);
    c.perm.assign(size_t(end - begin), 0);
    for (int64_t p = begin; p < end; ++p) c.perm[size_t(fill[size_t(c.cellOf[size_t(p - begin)])]++)] = p;
}

OCells* or_cells_build(const OAtoms* a,
                       int64_t begin,
                       int64_t end,
                       const double* delta,
                       const double* gmin,
                       const double* gmax)
{
    auto* c = new OCells();
    c->grid.init(gmin, gmax, delta);
    buildCells(*c, a->a, begin, end);
    return c;
}
void or_cells_free(OCells* c) { delete c; }
void or_cells_dims(const OCells* c, int* n) { for (int d = 0; d < 3; ++d) n[d] = c->grid.n[d]; }
void or_cells_get(const OCells* c, int* cellOf, int* counts, int64_t* offsets, int64_t* perm)
{
    if (cellOf) std::copy(c->cellOf.begin(), c->cellOf.end(), cellOf);
    if (counts) std::copy(c->counts.begin(), c->counts.end(), counts);
    if (offsets) std::copy(c->offsets.begin(), c->offsets.end() - 1, offsets);
    if (perm) std::copy(c->perm.begin(), c->perm.end(), perm);
}

// Atoms::permute (Atoms.hpp:95) -> Cabana::permute: new[begin+k] = old[perm[k]]
void or_atoms_permute(OAtoms* a, const OCells* c)
{
    std::vector<Atom> scratch(c->perm.size());
    for (size_t k = 0; k < c->perm.size(); ++k) scratch[k] = a->a[size_t(c->perm[k])];
    for (size_t k = 0; k < c->perm.size(); ++k) a->a[size_t(c->begin) + k] = scratch[k];
}

// ---- half Verlet list (Cabana VerletList<HalfNeighborTag, VerletLayout2D, TeamOpTag>) ----
struct OVerlet
{
    int64_t nrows = 0;  ///< slice size (counts exist for every particle, rows only in range)
    int64_t begin = 0, end = 0;
    int64_t maxN = 0;
    std::vector<int> counts;
    std::vector<int> neigh;  ///< row-major [row * maxN + n]
    Grid grid;
    int range = 1;
};

static void verletBuild(OVerlet& v,
                        const std::vector<Atom>& a,
                        int64_t nall,
                        int64_t begin,
                        int64_t end,
                        double rn,
                        double ratio,
                        const double* gmin,
                        const double* gmax,
                        int64_t maxNeigh,
                        bool half)
{
    v.nrows = nall;
    v.begin = begin;
    v.end = end;
    const double rsqr = rn * rn;
    const double gs = ratio * rn;
    const double delta[3] = {gs, gs, gs};
    v.grid.init(gmin, gmax, delta);
    v.range = int(std::ceil(1.0 / ratio));
    OCells cells;
    cells.grid = v.grid;
    buildCells(cells, a, 0, nall);
    const Grid& G = v.grid;
    auto row = [&](int64_t p, int* out, int64_t cap) -> int
    {
        const double* xp = a[size_t(p)].pos;
        const int cell = cells.cellOf[size_t(p)];
        const int ci = cell / (G.n[1] * G.n[2]);
        const int cj = (cell / G.n[2]) % G.n[1];
        const int ck = cell % G.n[2];
        const int imin = std::max(ci - v.range, 0), imax = std::min(ci + v.range + 1, G.n[0]);
        const int jmin = std::max(cj - v.range, 0), jmax = std::min(cj + v.range + 1, G.n[1]);
        const int kmin = std::max(ck - v.range, 0), kmax = std::min(ck + v.range + 1, G.n[2]);
        int cnt = 0;
        for (int i = imin; i < imax; ++i)
            for (int j = jmin; j < jmax; ++j)
                for (int k = kmin; k < kmax; ++k)
                {
                    if (G.minDistSqr(xp, i, j, k) > rsqr) continue;
                    const int c = G.cardinal(i, j, k);
                    for (int64_t s = cells.offsets[size_t(c)]; s < cells.offsets[size_t(c) + 1]; ++s)
                    {
                        const int64_t n = cells.perm[size_t(s)];
                        if (n == p) continue;
                        const double* xn = a[size_t(n)].pos;
                        if (half)
                        {
                            // Cabana NeighborDiscriminator<HalfNeighborTag>::isValid
                            bool valid = (xn[0] > xp[0]) ||
                                         ((xn[0] == xp[0]) &&
                                          ((xn[1] > xp[1]) || ((xn[1] == xp[1]) && (xn[2] > xp[2]))));
                            if (!valid) continue;
                        }
                        const double dx = xp[0] - xn[0];
                        const double dy = xp[1] - xn[1];
                        const double dz = xp[2] - xn[2];
                        const double d2 = dx * dx + dy * dy + dz * dz;
                        if (d2 <= rsqr)
                        {
                            if (cnt < cap) out[cnt] = int(n);
                            ++cnt;
                        }
                    }
                }
        return cnt;
    };
    v.counts.assign(size_t(nall), 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t p = begin; p < end; ++p) v.counts[size_t(p)] = row(p, nullptr, 0);
    int64_t mx = 0;
    for (int64_t p = begin; p < end; ++p) mx = std::max<int64_t>(mx, v.counts[size_t(p)]);
    v.maxN = std::max(mx, maxNeigh);
    v.neigh.assign(size_t(nall * v.maxN), -1);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t p = begin; p < end; ++p) row(p, &v.neigh[size_t(p * v.maxN)], v.maxN);
}

OVerlet* or_verlet_new() { return new OVerlet(); }
void or_verlet_free(OVerlet* v) { delete v; }
/// half != 0: Cabana half list; half == 0: the full list (both directions) for tests
void or_verlet_build(OVerlet* v,
                     const OAtoms* a,
                     int64_t nall,
                     int64_t begin,
                     int64_t end,
                     double rn,
                     double ratio,
                     const double* gmin,
                     const double* gmax,
                     int64_t maxNeigh,
                     int half)
{
    verletBuild(*v, a->a, nall, begin, end, rn, ratio, gmin, gmax, maxNeigh, half != 0);
}
void or_verlet_build_mol(OVerlet* v,
                         const OMolecules* m,
                         int64_t nall,
                         int64_t begin,
                         int64_t end,
                         double rn,
                         double ratio,
                         const double* gmin,
                         const double* gmax,
                         int64_t maxNeigh,
                         int half)
{
    std::vector<Atom> tmp(m->m.size());
    for (size_t i = 0; i < m->m.size(); ++i)
        for (int d = 0; d < 3; ++d) tmp[i].pos[d] = m->m[i].pos[d];
    verletBuild(*v, tmp, nall, begin, end, rn, ratio, gmin, gmax, maxNeigh, half != 0);
}
int64_t or_verlet_maxn(const OVerlet* v) { return v->maxN; }
int64_t or_verlet_nrows(const OVerlet* v) { return v->nrows; }
void or_verlet_counts(const OVerlet* v, int* out) { std::copy(v->counts.begin(), v->counts.end(), out); }
void or_verlet_neighbors(const OVerlet* v, int* out)
{
    std::copy(v->neigh.begin(), v->neigh.end(), out);
}
int64_t or_verlet_total(const OVerlet* v)
{
    int64_t s = 0;
    for (auto c : v->counts) s += c;
    return s;
}
void or_verlet_grid(const OVerlet* v, int* n, double* dx)
{
    for (int d = 0; d < 3; ++d)
    {
        n[d] = v->grid.n[d];
        dx[d] = v->grid.dx[d];
    }
}

// ---- Lennard-Jones ---------------------------------------------------------------------
struct OLJ
{
    std::vector<Pre> table;
    int64_t numTypes = 1;
    int64_t typeStride = 1;  ///< LennardJones uses 1 (LennardJones.cpp:52), LJ_IdealGas numTypes
    double rcSqr = 0;
    double energy = 0, virial = 0;
};

// LennardJones.cpp:81-116 and the init functor LennardJones.cpp:61-79
OLJ* or_lj_new(const double* cap,
               const double* rc,
               const double* sigma,
               const double* eps,
               int64_t numTypes,
               int shifted,
               int64_t typeStride)
{
    auto* lj = new OLJ();
    lj->numTypes = numTypes;
    lj->typeStride = typeStride;
    const int64_t nt2 = numTypes * numTypes;
    lj->table.assign(size_t(nt2), Pre{});
    double rcMax = rc[0];
    for (int64_t t = 0; t < nt2; ++t)
    {
        auto& p = lj->table[size_t(t)];
        auto sig2 = sigma[t] * sigma[t];
        auto sig6 = sig2 * sig2 * sig2;
        p.ff1 = 48.0 * eps[t] * sig6 * sig6;
        p.ff2 = 24.0 * eps[t] * sig6;
        p.ef1 = 4.0 * eps[t] * sig6 * sig6;
        p.ef2 = 4.0 * eps[t] * sig6;
        p.rcSqr = rc[t] * rc[t];
        p.capDist = cap[t];
        rcMax = std::max(rcMax, rc[t]);
    }
    for (int64_t t = 0; t < nt2; ++t)
    {
        auto& p = lj->table[size_t(t)];
        auto capDist = p.capDist;
        p.capDist = 0.0;
        p.capDistSqr = 0.0;
        auto fe = computeForceAndEnergy(p, capDist * capDist);
        p.capCoeff = fe.ff * capDist;
        p.eCap = fe.e;
        p.capDist = capDist;
        p.capDistSqr = capDist * capDist;
        if (shifted) p.shift = computeForceAndEnergy(p, p.rcSqr).e;
    }
    lj->rcSqr = rcMax * rcMax;
    return lj;
}
void or_lj_free(OLJ* lj) { delete lj; }
void or_lj_table(const OLJ* lj, double* out)
{
    std::memcpy(out, lj->table.data(), lj->table.size() * sizeof(Pre));
}
void or_lj_eval(const OLJ* lj, double distSqr, int64_t typeIdx, double* ff, double* e)
{
    auto r = computeForceAndEnergy(lj->table[size_t(typeIdx)], distSqr);
    *ff = r.ff;
    *e = r.e;
}

// LennardJones.hpp:136-206 (apply = apply_if with an always-true predicate)
void or_lj_apply(OLJ* lj, OAtoms* at, const OVerlet* v, const or_pred2* pred)
{
    auto& a = at->a;
    double E = 0.0, W = 0.0;
    const double rcSqr = lj->rcSqr;
#pragma omp parallel for reduction(+ : E, W) schedule(dynamic, 512)
    for (int64_t i = 0; i < at->nlocal; ++i)
    {
        const double xi[3] = {a[size_t(i)].pos[0], a[size_t(i)].pos[1], a[size_t(i)].pos[2]};
        double f[3] = {0.0, 0.0, 0.0};
        const int nn = v->counts[size_t(i)];
        for (int n = 0; n < nn; ++n)
        {
            const int64_t j = v->neigh[size_t(i * v->maxN + n)];
            const double* xj = a[size_t(j)].pos;
            if (pred && !evalPred2(*pred, xi, xj)) continue;
            const double dx = xi[0] - xj[0];
            const double dy = xi[1] - xj[1];
            const double dz = xi[2] - xj[2];
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > rcSqr) continue;
            const int64_t tIdx = a[size_t(i)].type * lj->typeStride + a[size_t(j)].type;
            auto fe = computeForceAndEnergy(lj->table[size_t(tIdx)], d2);
            E += fe.e;
            W -= 0.5 * fe.ff * d2;
            f[0] += dx * fe.ff;
            f[1] += dy * fe.ff;
            f[2] += dz * fe.ff;
            atomicAdd(a[size_t(j)].force[0], -(dx * fe.ff));
            atomicAdd(a[size_t(j)].force[1], -(dy * fe.ff));
            atomicAdd(a[size_t(j)].force[2], -(dz * fe.ff));
        }
        atomicAdd(a[size_t(i)].force[0], f[0]);
        atomicAdd(a[size_t(i)].force[1], f[1]);
        atomicAdd(a[size_t(i)].force[2], f[2]);
    }
    lj->energy = E;
    lj->virial = W;
}
double or_lj_energy(const OLJ* lj) { return lj->energy; }
double or_lj_virial(const OLJ* lj) { return lj->virial; }

// ---- integrators -----------------------------------------------------------------------
// VelocityVerlet.cpp:26-67 with UpdateSteps.hpp:25-54
double or_vv_pre(OAtoms* at, double dt)
{
    auto dtHalf = 0.5 * dt;
    double mx = 0.0;
#pragma omp parallel for reduction(max : mx)
    for (int64_t i = 0; i < at->nlocal; ++i)
    {
        auto& a = at->a[size_t(i)];
        double old[3] = {a.pos[0], a.pos[1], a.pos[2]};
        auto dtfm = dtHalf / a.mass;
        for (int d = 0; d < 3; ++d) a.vel[d] += dtfm * a.force[d];
        for (int d = 0; d < 3; ++d) a.pos[d] += dt * a.vel[d];
        double dx[3] = {old[0] - a.pos[0], old[1] - a.pos[1], old[2] - a.pos[2]};
        double d2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        mx = std::max(d2, mx);
    }
    return std::sqrt(mx);
}

// VelocityVerlet.cpp:69-90
void or_vv_post(OAtoms* at, double dt)
{
    auto dtHalf = 0.5 * dt;
#pragma omp parallel for
    for (int64_t i = 0; i < at->nlocal; ++i)
    {
        auto& a = at->a[size_t(i)];
        auto dtfm = dtHalf / a.mass;
        for (int d = 0; d < 3; ++d) a.vel[d] += dtfm * a.force[d];
    }
}

// VelocityVerletLangevinThermostat.hpp:63-136 with UpdateSteps.hpp:56-79; the noise for
// atom i at call c is normals3(seed, i, c)
double or_langevin_pre(OAtoms* at,
                       double dt,
                       double zeta,
                       double temperature,
                       uint64_t seed,
                       uint64_t call,
                       const or_pred1* pred)
{
    auto dtHalf = 0.5 * dt;
    auto dtFull = dt;
    double mx = 0.0;
#pragma omp parallel for reduction(max : mx)
    for (int64_t i = 0; i < at->nlocal; ++i)
    {
        auto& a = at->a[size_t(i)];
        double old[3] = {a.pos[0], a.pos[1], a.pos[2]};
        auto dtfm = dtHalf / a.mass;
        for (int d = 0; d < 3; ++d) a.vel[d] += dtfm * a.force[d];
        for (int d = 0; d < 3; ++d) a.pos[d] += dtHalf * a.vel[d];
        if (!pred || evalPred1(*pred, a.pos[0], a.pos[1], a.pos[2]))
        {
            double xi[3];
            normals3(seed, uint64_t(i), call, xi);
            auto dtm = dtFull / a.mass;
            auto damping = detExp(-zeta * dtm);
            auto sigma = std::sqrt(temperature / a.mass * (1.0 - detExp(-2.0 * zeta * dtm)));
            for (int d = 0; d < 3; ++d) a.vel[d] *= damping;
            for (int d = 0; d < 3; ++d) a.vel[d] += sigma * xi[d];
        }
        for (int d = 0; d < 3; ++d) a.pos[d] += dtHalf * a.vel[d];
        double dx[3] = {old[0] - a.pos[0], old[1] - a.pos[1], old[2] - a.pos[2]};
        double d2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        mx = std::max(d2, mx);
    }
    return std::sqrt(mx);
}

// analysis/KineticEnergy.hpp:26-39
double or_kinetic_energy(const OAtoms* at)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s)
    for (int64_t i = 0; i < at->nlocal; ++i)
    {
        const auto& a = at->a[size_t(i)];
        s += a.mass * (a.vel[0] * a.vel[0] + a.vel[1] * a.vel[1] + a.vel[2] * a.vel[2]);
    }
    return 0.5 * s;
}

// ---- AdResS: weighting, molecules ------------------------------------------------------
// util/math.hpp:31-46
static double powInt(double x, int64_t n)
{
    auto ww = x;
    auto yy = 1.0;
    for (int64_t nn = (n > 0) ? n : -n; nn != 0; nn >>= 1)
    {
        if ((nn & 1) == 1) yy *= ww;
        ww *= ww;
    }
    return (n > 0) ? yy : 1.0 / yy;
}

// weighting_function/Slab.hpp:96-187
static void slabEval(const or_slab& s, double x, double* lambda, double* mod, double* g)
{
    const double halfAT = 0.5 * s.atomistic_diameter;
    const double hy = s.hybrid_diameter;
    const int64_t expo = 2 * s.nu;
    double dx = x - s.center[0];
    double absDx = std::abs(dx);
    g[1] = 0.0;
    g[2] = 0.0;
    if (absDx < halfAT || (!(absDx > halfAT + hy) && s.interface_type == 1))
    {
        *lambda = 1.0;
        *mod = 1.0;
        g[0] = 0.0;
        return;
    }
    if (absDx > halfAT + hy)
    {
        *lambda = 0.0;
        *mod = 0.0;
        g[0] = 0.0;
        return;
    }
    const double arg = PI / (2.0 * hy) * (absDx - halfAT);
    auto base = std::cos(arg);
    *lambda = base * base;
    *mod = powInt(base, expo);
    auto factor = -PI / (2.0 * hy) * double(expo) * std::sin(arg) * powInt(base, expo - 1) / absDx;
    g[0] = factor * dx;
}

void or_slab_eval(const or_slab* s, double x, double* out5)
{
    double g[3];
    slabEval(*s, x, &out5[0], &out5[1], g);
    out5[2] = g[0];
    out5[3] = g[1];
    out5[4] = g[2];
}

// weighting_function/Spherical.hpp:37-76 (center, atomistic radius, hybrid shell width,
// exponent); out4 = lambda, gradient x, y, z
void or_spherical_eval(const double* center, double atomisticRadius, double hybridDiameter, int64_t exponent,
                       double x, double y, double z, double* out4)
{
    const double dx[3] = {x - center[0], y - center[1], z - center[2]};
    const double dxSqr = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    const double atSqr = atomisticRadius * atomisticRadius;
    const double cgSqr = (atomisticRadius + hybridDiameter) * (atomisticRadius + hybridDiameter);
    if (dxSqr < atSqr || dxSqr > cgSqr)
    {
        out4[0] = dxSqr < atSqr ? 1.0 : 0.0;
        out4[1] = out4[2] = out4[3] = 0.0;
        return;
    }
    const double r = std::sqrt(dxSqr);
    const double arg = PI / (2.0 * hybridDiameter) * (r - atomisticRadius);
    const double base = std::cos(arg);
    out4[0] = powInt(base, exponent);
    const double factor =
        -PI / (2.0 * hybridDiameter) * double(exponent) * std::sin(arg) * powInt(base, exponent - 1) / r;
    for (int d = 0; d < 3; ++d) out4[1 + d] = factor * dx[d];
}

// action/BerendsenThermostat.cpp:27-50; returns 0, or -1 for a non-positive target
int or_berendsen_apply(OAtoms* at, double current, double target, double gamma)
{
    if (current <= 0.0) return 0;
    if (!(target > 0.0)) return -1;
    const double beta = std::sqrt(1.0 + gamma * (target / current - 1.0));
#pragma omp parallel for
    for (int64_t i = 0; i < at->nlocal; ++i)
        for (int d = 0; d < 3; ++d) at->a[size_t(i)].vel[d] *= beta;
    return 0;
}

// action/UpdateMolecules.hpp:24-70
void or_update_molecules(OMolecules* m, const OAtoms* at, const or_slab* s)
{
#pragma omp parallel for
    for (int64_t mi = 0; mi < m->nlocal + m->nghost; ++mi)
    {
        auto& mo = m->m[size_t(mi)];
        mo.pos[0] = mo.pos[1] = mo.pos[2] = 0.0;
        for (int64_t ai = mo.offset; ai < mo.offset + mo.numAtoms; ++ai)
            for (int d = 0; d < 3; ++d)
                mo.pos[d] += at->a[size_t(ai)].pos[d] * at->a[size_t(ai)].relMass;
        if (s) slabEval(*s, mo.pos[0], &mo.lambda, &mo.modLambda, mo.grad);
    }
}

// action/ContributeMoleculeForceToAtoms.cpp:23-48
void or_contribute_molecule_force(const OMolecules* m, OAtoms* at)
{
#pragma omp parallel for
    for (int64_t mi = 0; mi < m->nlocal + m->nghost; ++mi)
    {
        const auto& mo = m->m[size_t(mi)];
        for (int64_t ai = mo.offset; ai < mo.offset + mo.numAtoms; ++ai)
        {
            auto& a = at->a[size_t(ai)];
            for (int d = 0; d < 3; ++d) atomicAdd(a.force[d], a.relMass * mo.force[d]);
        }
    }
}

// MultiResRealAtomsExchange.cpp:23-73
void or_multires_exchange_real(const or_box* box, OMolecules* m, OAtoms* at)
{
    auto s = makeSub(*box);
#pragma omp parallel for
    for (int64_t mi = 0; mi < m->nlocal; ++mi)
    {
        auto& mo = m->m[size_t(mi)];
        for (int d = 0; d < 3; ++d)
        {
            auto& X = mo.pos[d];
            if (s.max[d] <= X)
            {
                X -= s.diam[d];
                for (int64_t ai = mo.offset; ai < mo.offset + mo.numAtoms; ++ai)
                    at->a[size_t(ai)].pos[d] -= s.diam[d];
            }
            if (X < s.min[d])
            {
                X += s.diam[d];
                for (int64_t ai = mo.offset; ai < mo.offset + mo.numAtoms; ++ai)
                    at->a[size_t(ai)].pos[d] += s.diam[d];
            }
        }
    }
}

// MultiResPeriodicGhostExchange.cpp:60-248: molecules with pos >= maxInner are copied
// shifted by -L first, then molecules with pos < minInner shifted by +L.
void or_multires_ghost_axis(OGhost* g, OMolecules* m, OAtoms* at, const or_box* box, int axis)
{
    auto s = makeSub(*box);
    const int64_t totM = m->nlocal + m->nghost;
    const int64_t totA = at->nlocal + at->nghost;
    std::vector<int64_t> P, PA, N, NA;  // positive: >= maxInner, negative: < minInner
    int64_t pa = 0, na = 0;
    for (int64_t i = 0; i < totM; ++i)
    {
        const double x = m->m[size_t(i)].pos[axis];
        if (x < s.minInner[axis])
        {
            N.push_back(i);
            NA.push_back(na);
            na += m->m[size_t(i)].numAtoms;
        }
        if (x >= s.maxInner[axis])
        {
            P.push_back(i);
            PA.push_back(pa);
            pa += m->m[size_t(i)].numAtoms;
        }
    }
    const int64_t nP = int64_t(P.size()), nN = int64_t(N.size());
    m->m.resize(size_t(totM + nP + nN));
    at->a.resize(size_t(totA + pa + na));
    if (int64_t(g->corr.size()) < totA + pa + na) g->corr.resize(size_t(totA + pa + na), 0);
    auto copyMol = [&](int64_t src, int64_t dstM, int64_t dstA, double shift)
    {
        const auto srcOff = m->m[size_t(src)].offset;
        const auto size = m->m[size_t(src)].numAtoms;
        m->m[size_t(dstM)] = m->m[size_t(src)];
        m->m[size_t(dstM)].pos[axis] += shift;
        m->m[size_t(dstM)].offset = dstA;
        m->m[size_t(dstM)].numAtoms = size;
        for (int64_t k = 0; k < size; ++k)
        {
            at->a[size_t(dstA + k)] = at->a[size_t(srcOff + k)];
            at->a[size_t(dstA + k)].pos[axis] += shift;
            auto real = srcOff + k;
            while (g->corr[size_t(real)] != -1) real = g->corr[size_t(real)];
            g->corr[size_t(dstA + k)] = real;
        }
    };
    for (int64_t k = 0; k < nP; ++k) copyMol(P[size_t(k)], totM + k, totA + PA[size_t(k)], -s.diam[axis]);
    for (int64_t k = 0; k < nN; ++k)
        copyMol(N[size_t(k)], totM + nP + k, totA + pa + NA[size_t(k)], +s.diam[axis]);
    m->nghost += nP + nN;
    at->nghost += pa + na;
}

// MultiResPeriodicGhostExchange.cpp:250-265
void or_multires_ghost_xyz(OGhost* g, OMolecules* m, OAtoms* at, const or_box* box)
{
    m->nghost = 0;
    g->corr.assign(size_t(at->nlocal), -1);
    at->nghost = 0;
    for (int d = 0; d < 3; ++d) or_multires_ghost_axis(g, m, at, box, d);
}

// ---- LJ_IdealGas (action/LJ_IdealGas.cpp:21-292) ----------------------------------------
struct OLJIG
{
    OLJ* lj = nullptr;
    int64_t numTypes = 1;
    double rcSqr = 0;
    Hist comp, compCount, meanComp;
    int64_t runCounter = 0;
    int64_t samplingInterval = 200;
    int64_t updateInterval = 20000;
};

OLJIG* or_ljig_new(const double* cap,
                   const double* rc,
                   const double* sigma,
                   const double* eps,
                   int64_t numTypes,
                   int shifted)
{
    auto* s = new OLJIG();
    s->lj = or_lj_new(cap, rc, sigma, eps, numTypes, shifted, numTypes);
    s->numTypes = numTypes;
    s->rcSqr = s->lj->rcSqr;
    s->comp.init(0.0, 1.0, 200, numTypes);
    s->compCount.init(0.0, 1.0, 200, numTypes);
    s->meanComp.init(0.0, 1.0, 200, numTypes);
    return s;
}
void or_ljig_free(OLJIG* s)
{
    or_lj_free(s->lj);
    delete s;
}
void or_ljig_set_intervals(OLJIG* s, int64_t sampling, int64_t update)
{
    s->samplingInterval = sampling;
    s->updateInterval = update;
}
void or_ljig_get_hist(const OLJIG* s, int which, double* out)
{
    const Hist& h = which == 0 ? s->comp : (which == 1 ? s->compCount : s->meanComp);
    std::copy(h.data.begin(), h.data.end(), out);
}
void or_ljig_set_mean(OLJIG* s, const double* in)
{
    std::copy(in, in + s->meanComp.data.size(), s->meanComp.data.begin());
}

static inline bool isAT(double l) { return l >= 1.0; }
static inline bool isCG(double l) { return l <= 0.0; }
static inline bool isHY(double l) { return !isAT(l) && !isCG(l); }

// LJ_IdealGas.cpp:21-50
static void updateMeanComp(Hist& comp, Hist& cnt, Hist& mean, double factor)
{
    for (int64_t b = 0; b < comp.numBins; ++b)
        for (int64_t h = 0; h < comp.numHist; ++h)
        {
            if (cnt.at(b, h) < 0.5) continue;
            auto e = comp.at(b, h) / cnt.at(b, h);
            mean.at(b, h) = (factor * mean.at(b, h) + e) / (factor + 1.0);
            comp.at(b, h) = 0.0;
            cnt.at(b, h) = 0.0;
        }
}

// LJ_IdealGas.cpp:52-260
double or_ljig_run(OLJIG* s, OMolecules* mols, const OVerlet* v, OAtoms* atoms)
{
    auto& M = mols->m;
    auto& A = atoms->a;
    const bool sampling = (s->runCounter % s->samplingInterval) == 0;
    const int64_t nt = s->numTypes;
    double energy = 0.0;
#pragma omp parallel for reduction(+ : energy) schedule(dynamic, 64)
    for (int64_t alpha = 0; alpha < mols->nlocal; ++alpha)
    {
        double fAlpha[3] = {0.0, 0.0, 0.0};
        const auto mlA = M[size_t(alpha)].modLambda;
        int64_t binAlpha = -1;
        if (isHY(mlA)) binAlpha = s->comp.getBin(M[size_t(alpha)].lambda);
        const double* gA = M[size_t(alpha)].grad;
        const auto startA = M[size_t(alpha)].offset;
        const auto endA = startA + M[size_t(alpha)].numAtoms;
        const int nn = v->counts[size_t(alpha)];
        for (int n = 0; n < nn; ++n)
        {
            const int64_t beta = v->neigh[size_t(alpha * v->maxN + n)];
            double fBeta[3] = {0.0, 0.0, 0.0};
            const auto mlB = M[size_t(beta)].modLambda;
            const double* gB = M[size_t(beta)].grad;
            const auto w = 0.5 * (mlA + mlB);
            if (isCG(mlA) && isCG(mlB)) continue;
            const auto startB = M[size_t(beta)].offset;
            const auto endB = startB + M[size_t(beta)].numAtoms;
            for (int64_t i = startA; i < endA; ++i)
            {
                const double xi[3] = {A[size_t(i)].pos[0], A[size_t(i)].pos[1], A[size_t(i)].pos[2]};
                double fi[3] = {0.0, 0.0, 0.0};
                for (int64_t j = startB; j < endB; ++j)
                {
                    const double dx = xi[0] - A[size_t(j)].pos[0];
                    const double dy = xi[1] - A[size_t(j)].pos[1];
                    const double dz = xi[2] - A[size_t(j)].pos[2];
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 > s->rcSqr) continue;
                    auto tIdx = A[size_t(i)].type * nt + A[size_t(j)].type;
                    auto fe = computeForceAndEnergy(s->lj->table[size_t(tIdx)], d2);
                    auto ff = fe.ff * w;
                    fi[0] += dx * ff;
                    fi[1] += dy * ff;
                    fi[2] += dz * ff;
                    atomicAdd(A[size_t(j)].force[0], -(dx * ff));
                    atomicAdd(A[size_t(j)].force[1], -(dy * ff));
                    atomicAdd(A[size_t(j)].force[2], -(dz * ff));
                    energy += fe.e * w;
                    auto Vij = 0.5 * fe.e;
                    if (isHY(mlA) || isHY(mlB))
                    {
                        for (int d = 0; d < 3; ++d) fAlpha[d] += -Vij * gA[d];
                        for (int d = 0; d < 3; ++d) fBeta[d] += -Vij * gB[d];
                        if (sampling)
                        {
                            int64_t binBeta = -1;
                            if (isHY(mlB)) binBeta = s->comp.getBin(M[size_t(beta)].lambda);
                            if (isHY(mlA) && binAlpha != -1)
                                atomicAdd(s->comp.at(binAlpha, A[size_t(i)].type), Vij);
                            if (isHY(mlB) && binBeta != -1)
                                atomicAdd(s->comp.at(binBeta, A[size_t(j)].type), Vij);
                        }
                    }
                }
                for (int d = 0; d < 3; ++d) atomicAdd(A[size_t(i)].force[d], fi[d]);
            }
            for (int d = 0; d < 3; ++d) atomicAdd(M[size_t(beta)].force[d], fBeta[d]);
        }
        if (sampling)
            for (int64_t ai = startA; ai < endA; ++ai)
                if (isHY(mlA) && binAlpha != -1)
                    atomicAdd(s->compCount.at(binAlpha, A[size_t(ai)].type), 1.0);
        if (isHY(mlA) && binAlpha != -1)
        {
            const auto mc = s->meanComp.at(binAlpha, A[size_t(startA)].type);
            for (int d = 0; d < 3; ++d) fAlpha[d] += mc * gA[d];
        }
        for (int d = 0; d < 3; ++d) atomicAdd(M[size_t(alpha)].force[d], fAlpha[d]);
    }
    if (s->runCounter % s->updateInterval == 0) updateMeanComp(s->comp, s->compCount, s->meanComp, 10.0);
    ++s->runCounter;
    return energy;
}

// ---- thermodynamic force (action/ThermodynamicForce.cpp:25-130, .hpp:98-216) ----------
struct OTF
{
    Hist force, density;
    int64_t samples = 0;
    double binVolume = 1;
    std::vector<double> forceFactor;
    bool sym = false, periodic = false;
};

OTF* or_tf_new(const double* target,
               const double* modulation,
               int64_t numTypes,
               const or_box* box,
               double binWidth,
               int sym,
               int periodic)
{
    auto s = makeSub(*box);
    auto* t = new OTF();
    const auto nb = int64_t(std::ceil(s.diam[0] / binWidth));
    t->force.init(s.min[0], s.max[0], nb, numTypes);
    t->density.init(s.min[0], s.max[0], nb, numTypes);
    t->binVolume = s.diam[1] * s.diam[2] * t->density.binSize;
    for (int64_t i = 0; i < numTypes; ++i) t->forceFactor.push_back(modulation[i] / target[i]);
    t->sym = sym != 0;
    t->periodic = periodic != 0;
    return t;
}
void or_tf_free(OTF* t) { delete t; }
int64_t or_tf_num_bins(const OTF* t) { return t->force.numBins; }
int64_t or_tf_samples(const OTF* t) { return t->samples; }
void or_tf_get(const OTF* t, int which, double* out)
{
    const Hist& h = which == 0 ? t->force : t->density;
    std::copy(h.data.begin(), h.data.end(), out);
}
void or_tf_set_force(OTF* t, const double* in)
{
    std::copy(in, in + t->force.data.size(), t->force.data.begin());
}

// ThermodynamicForce.cpp:74-86 with analysis/AxialDensityProfile.cpp:21-51
void or_tf_sample(OTF* t, const OAtoms* at)
{
    Hist h;
    h.init(t->density.min, t->density.max, t->density.numBins, t->density.numHist);
    for (int64_t i = 0; i < at->nlocal; ++i)
    {
        auto bin = h.getBin(at->a[size_t(i)].pos[0]);
        if (bin == -1) continue;
        h.at(bin, at->a[size_t(i)].type) += 1.0;
    }
    for (size_t k = 0; k < h.data.size(); ++k) t->density.data[k] += h.data[k];
    ++t->samples;
}

// ThermodynamicForce.hpp:124-152 (update = update_if with an always-true predicate)
int or_tf_update(OTF* t, double sigma, double range, const or_pred1* pred)
{
    if (t->samples <= 0) return -1;
    if (t->sym) t->density.makeSymmetric();
    auto norm = 1.0 / (t->binVolume * double(t->samples));
    t->density.scale(norm);
    auto sm = smoothen(t->density, sigma, range, t->periodic);
    auto gr = gradient(sm, t->periodic);
    for (int64_t b = 0; b < gr.numBins; ++b)
        for (int64_t h = 0; h < gr.numHist; ++h) gr.at(b, h) *= t->forceFactor[size_t(h)];
    for (int64_t b = 0; b < gr.numBins; ++b)
    {
        const double x = gr.getBinPosition(b);
        if (pred && !evalPred1(*pred, x, x, x))
            for (int64_t h = 0; h < gr.numHist; ++h) gr.at(b, h) = 0.0;
    }
    for (size_t k = 0; k < t->force.data.size(); ++k) t->force.data[k] -= gr.data[k];
    std::fill(t->density.data.begin(), t->density.data.end(), 0.0);
    t->samples = 0;
    return 0;
}

// ThermodynamicForce.hpp:98-122
void or_tf_apply(const OTF* t, OAtoms* at, const or_pred1* pred)
{
#pragma omp parallel for
    for (int64_t i = 0; i < at->nlocal; ++i)
    {
        auto& a = at->a[size_t(i)];
        if (pred && !evalPred1(*pred, a.pos[0], a.pos[1], a.pos[2])) continue;
        auto bin = t->force.getBin(a.pos[0]);
        if (bin != -1) a.force[0] += t->force.at(bin, a.type);
    }
}

// ThermodynamicForce.hpp:154-216 with util/interpolation.hpp:34-38
void or_tf_apply_interpolated(const OTF* t, OAtoms* at, const or_pred1* pred)
{
    const Hist& F = t->force;
#pragma omp parallel for
    for (int64_t i = 0; i < at->nlocal; ++i)
    {
        auto& a = at->a[size_t(i)];
        auto x = a.pos[0];
        if (pred && !evalPred1(*pred, a.pos[0], a.pos[1], a.pos[2])) continue;
        auto bin = F.getBin(x);
        if (bin == -1) continue;
        auto binStart = F.min + double(bin) * F.binSize;
        auto frac = (x - binStart) * F.invBinSize;
        int64_t l, r;
        double fac;
        if (frac < 0.5)
        {
            l = bin - 1;
            r = bin;
            fac = frac + 0.5;
        }
        else
        {
            l = bin;
            r = bin + 1;
            fac = frac - 0.5;
        }
        if (l >= 0 && r < F.numBins)
        {
            auto L = F.at(l, a.type), R = F.at(r, a.type);
            a.force[0] += L + (R - L) * fac;
        }
        else
            a.force[0] += F.at(bin, a.type);
    }
}

// ThermodynamicForce.cpp:98-130
void or_tf_mu(const OTF* t, double* left, double* right)
{
    const Hist& F = t->force;
    for (int64_t h = 0; h < F.numHist; ++h)
    {
        double l = 0.0, r = 0.0;
        for (int64_t b = 0; b < F.numBins / 2; ++b) l += F.at(b, h);
        for (int64_t b = F.numBins / 2; b < F.numBins; ++b) r += F.at(b, h);
        left[h] = l * F.binSize;
        right[h] = r * F.binSize;
    }
}

// ---- MultiHistogram ops exposed for the histogram golden tests -------------------------
// in/out arrays are [numBins * numHist]
void or_hist_gradient(double mn, double mx, int64_t nb, int64_t nh, const double* in, int periodic, double* out)
{
    Hist h;
    h.init(mn, mx, nb, nh);
    std::copy(in, in + nb * nh, h.data.begin());
    auto g = gradient(h, periodic != 0);
    std::copy(g.data.begin(), g.data.end(), out);
}
void or_hist_smoothen(double mn,
                      double mx,
                      int64_t nb,
                      int64_t nh,
                      const double* in,
                      double sigma,
                      double range,
                      int periodic,
                      double* out)
{
    Hist h;
    h.init(mn, mx, nb, nh);
    std::copy(in, in + nb * nh, h.data.begin());
    auto g = smoothen(h, sigma, range, periodic != 0);
    std::copy(g.data.begin(), g.data.end(), out);
}
void or_hist_make_symmetric(int64_t nb, int64_t nh, double* data)
{
    Hist h;
    h.init(0.0, 1.0, nb, nh);
    std::copy(data, data + nb * nh, h.data.begin());
    h.makeSymmetric();
    std::copy(h.data.begin(), h.data.end(), data);
}
int64_t or_hist_get_bin(double mn, double mx, int64_t nb, double v)
{
    Hist h;
    h.init(mn, mx, nb, 1);
    return h.getBin(v);
}
double or_hist_bin_position(double mn, double mx, int64_t nb, int64_t b)
{
    Hist h;
    h.init(mn, mx, nb, 1);
    return h.getBinPosition(b);
}

// ---- the reference's time loop (examples/02_LennardJones_NVE/02_LennardJones_NVE.cpp:
// 135-186; with langevin != 0 the pre-force integrator of examples/01; with permute != 0
// the LinkedCellList + Atoms::permute of tests/NVT/NVT.cpp:138-144 on rebuild).
// energies_out[3*r] = (PE, KE, virial) after the force evaluation of step r*energy_every.
int64_t or_run_md(OAtoms* at, const or_box* box, const or_md_params* prm, double* energies_out)
{
    auto s = makeSub(*box);
    const double rn = prm->rc + prm->skin;
    const double cap = prm->cap, rc = prm->rc, sig = prm->sigma, eps = prm->epsilon;
    OLJ* lj = or_lj_new(&cap, &rc, &sig, &eps, 1, prm->shifted, 1);
    OGhost ghost;
    OVerlet vl;
    double maxDisp = std::numeric_limits<double>::max();
    int64_t rebuilds = 0, rec = 0;
    for (int64_t step = 0; step < prm->nsteps; ++step)
    {
        if (prm->langevin)
            maxDisp += or_langevin_pre(at, prm->dt, prm->gamma, prm->temperature, prm->seed,
                                       uint64_t(step), nullptr);
        else
            maxDisp += or_vv_pre(at, prm->dt);
        if (maxDisp >= prm->skin * 0.5)
        {
            maxDisp = 0.0;
            or_map_into_domain(at, box);
            if (prm->permute)
            {
                const double delta[3] = {rn, rn, rn};
                OCells* cells = or_cells_build(at, 0, at->nlocal, delta, s.min, s.max);
                or_atoms_permute(at, cells);
                or_cells_free(cells);
            }
            or_ghost_create_xyz(&ghost, at, box);
            verletBuild(vl, at->a, at->nlocal + at->nghost, 0, at->nlocal, rn, prm->cell_ratio,
                        s.minGhost, s.maxGhost, prm->max_neigh, true);
            ++rebuilds;
        }
        else
        {
            or_ghost_update(&ghost, at, box);
        }
        for (int64_t i = 0; i < at->nlocal + at->nghost; ++i)
            for (int d = 0; d < 3; ++d) at->a[size_t(i)].force[d] = 0.0;
        or_lj_apply(lj, at, &vl, nullptr);
        or_ghost_fold(&ghost, at);
        or_vv_post(at, prm->dt);
        if (energies_out && prm->energy_every > 0 && step % prm->energy_every == 0)
        {
            energies_out[3 * rec + 0] = lj->energy;
            energies_out[3 * rec + 1] = or_kinetic_energy(at);
            energies_out[3 * rec + 2] = lj->virial;
            ++rec;
        }
    }
    or_lj_free(lj);
    return rebuilds;
}

// ---- adjoint of the LJ force map (new code, no reference counterpart) ------------------
// L = sum_i g_i . F_i.  Each half-list pair (i, j) with j' the real atom of j contributes
//   dL/dx_i += J u,  dL/dx_j' -= J u,  u = g_i - g_j',  J = s I + 2 s'(r^2) d d^T
//   dL/deps_t += u.d s/eps,  dL/dsigma_t += u.d ds/dsigma  (t = type_i*stride + type_j)
// and the parameter gradients are finally symmetrised over (a,b) and (b,a).
void or_lj_adjoint(const OLJ* lj,
                   const OAtoms* at,
                   const OVerlet* v,
                   const OGhost* g,
                   const double* dLdF,
                   double* dLdx,
                   double* dLdeps,
                   double* dLdsigma)
{
    const auto& a = at->a;
    const int64_t nl = at->nlocal;
    const int64_t nt2 = int64_t(lj->table.size());
    std::vector<double> ge(size_t(nt2), 0.0), gs(size_t(nt2), 0.0);
    for (int64_t k = 0; k < 3 * nl; ++k) dLdx[k] = 0.0;
    auto realOf = [&](int64_t j) -> int64_t
    {
        if (j < nl) return j;
        return g->corr[size_t(j)];
    };
    for (int64_t i = 0; i < nl; ++i)
    {
        const int nn = v->counts[size_t(i)];
        for (int n = 0; n < nn; ++n)
        {
            const int64_t j = v->neigh[size_t(i * v->maxN + n)];
            const int64_t jr = realOf(j);
            const double dx[3] = {a[size_t(i)].pos[0] - a[size_t(j)].pos[0], a[size_t(i)].pos[1] - a[size_t(j)].pos[1],
                                  a[size_t(i)].pos[2] - a[size_t(j)].pos[2]};
            const double d2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            if (d2 > lj->rcSqr) continue;
            const int64_t t = a[size_t(i)].type * lj->typeStride + a[size_t(j)].type;
            const Pre& p = lj->table[size_t(t)];
            const double sigma = std::pow(p.ef1 / p.ef2, 1.0 / 6.0);
            const double eps = p.ef2 * p.ef2 / (4.0 * p.ef1);
            double u[3], du = 0.0;
            for (int d = 0; d < 3; ++d)
            {
                u[d] = dLdF[3 * i + d] - dLdF[3 * jr + d];
                du += dx[d] * u[d];
            }
            double s, sp, dsds;
            if (d2 >= p.capDistSqr)
            {
                const double r2i = 1.0 / d2, r6i = r2i * r2i * r2i;
                s = r6i * (p.ff1 * r6i - p.ff2) * r2i;
                sp = -(7.0 * p.ff1 * r6i * r6i * r2i * r2i - 4.0 * p.ff2 * r6i * r2i * r2i);
                dsds = r6i * (12.0 * p.ff1 * r6i - 6.0 * p.ff2) * r2i / sigma;
            }
            else
            {
                const double r = std::sqrt(d2);
                s = p.capCoeff / r;
                sp = -0.5 * p.capCoeff / (r * d2);
                const double c2i = 1.0 / p.capDistSqr, c6i = c2i * c2i * c2i;
                dsds = (c6i * (12.0 * p.ff1 * c6i - 6.0 * p.ff2) * c2i / sigma) * p.capDist / r;
            }
            for (int d = 0; d < 3; ++d)
            {
                const double vd = s * u[d] + 2.0 * sp * du * dx[d];
                dLdx[3 * i + d] += vd;
                dLdx[3 * jr + d] -= vd;
            }
            ge[size_t(t)] += du * s / eps;
            gs[size_t(t)] += du * dsds;
        }
    }
    const int64_t n = lj->numTypes, st = lj->typeStride;
    for (int64_t k = 0; k < nt2; ++k) dLdeps[k] = dLdsigma[k] = 0.0;
    for (int64_t x = 0; x < n; ++x)
        for (int64_t y = 0; y < n; ++y)
        {
            const int64_t i1 = x * st + y, i2 = y * st + x;
            if (i1 >= nt2 || i2 >= nt2) continue;
            dLdeps[i1] = i1 == i2 ? ge[size_t(i1)] : ge[size_t(i1)] + ge[size_t(i2)];
            dLdsigma[i1] = i1 == i2 ? gs[size_t(i1)] : gs[size_t(i1)] + gs[size_t(i2)];
        }
}

// ---- stateful form of the same loop (for bounded CPU baseline samples) ----------------
struct OMD
{
    OAtoms* at = nullptr;
    or_box box{};
    or_md_params prm{};
    OLJ* lj = nullptr;
    OGhost ghost;
    OVerlet vl;
    double maxDisp = std::numeric_limits<double>::max();
    int64_t step = 0, rebuilds = 0;
};

OMD* or_md_new(OAtoms* at, const or_box* box, const or_md_params* prm)
{
    auto* md = new OMD();
    md->at = at;
    md->box = *box;
    md->prm = *prm;
    const double cap = prm->cap, rc = prm->rc, sig = prm->sigma, eps = prm->epsilon;
    md->lj = or_lj_new(&cap, &rc, &sig, &eps, 1, prm->shifted, 1);
    return md;
}

void or_md_free(OMD* md)
{
    or_lj_free(md->lj);
    delete md;
}

// runs n steps of the loop of or_run_md; returns the rebuild count so far
int64_t or_md_steps(OMD* md, int64_t n)
{
    auto* at = md->at;
    const auto* prm = &md->prm;
    auto s = makeSub(md->box);
    const double rn = prm->rc + prm->skin;
    for (int64_t k = 0; k < n; ++k, ++md->step)
    {
        if (prm->langevin)
            md->maxDisp += or_langevin_pre(at, prm->dt, prm->gamma, prm->temperature, prm->seed,
                                           uint64_t(md->step), nullptr);
        else
            md->maxDisp += or_vv_pre(at, prm->dt);
        if (md->maxDisp >= prm->skin * 0.5)
        {
            md->maxDisp = 0.0;
            or_map_into_domain(at, &md->box);
            if (prm->permute)
            {
                const double delta[3] = {rn, rn, rn};
                OCells* cells = or_cells_build(at, 0, at->nlocal, delta, s.min, s.max);
                or_atoms_permute(at, cells);
                or_cells_free(cells);
            }
            or_ghost_create_xyz(&md->ghost, at, &md->box);
            verletBuild(md->vl, at->a, at->nlocal + at->nghost, 0, at->nlocal, rn, prm->cell_ratio, s.minGhost,
                        s.maxGhost, prm->max_neigh, true);
            ++md->rebuilds;
        }
        else
            or_ghost_update(&md->ghost, at, &md->box);
#pragma omp parallel for
        for (int64_t i = 0; i < at->nlocal + at->nghost; ++i)
            for (int d = 0; d < 3; ++d) at->a[size_t(i)].force[d] = 0.0;
        or_lj_apply(md->lj, at, &md->vl, nullptr);
        or_ghost_fold(&md->ghost, at);
        or_vv_post(at, prm->dt);
    }
    return md->rebuilds;
}

double or_md_energy(const OMD* md) { return md->lj->energy; }

}  // extern "C"
