// Pair potential, weighting function and histogram arithmetic shared by the force,
// AdResS and ensemble kernels (device and host).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "mrmd_hip.h"

namespace mrmd_hip
{
constexpr double PI = 3.14159265358979323846;

// 1 / x from the hardware reciprocal estimate refined by one Newton-Raphson step
// (e = 1 - x r, r <- r + r e), which squares its relative error: measured on gfx950 over
// 16.8 M arguments in [1e-3, 1e3), the estimate is off by up to 2.6e8 ulp (~2^-24.5), the
// refined value by at most 11 ulp (2.4e-15 relative; a second step gives the correctly
// rounded quotient but costs 2 FMAs per pair).  The pair forces it feeds stay within a few
// 1e-14 of the correctly rounded arithmetic.  x = 0 gives a NaN.
__device__ inline double recipNewton(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// PrecomputedValues (LennardJones.hpp:27-39)
struct Pre
{
    double ff1, ff2, ef1, ef2, rcSqr, capDist, capDistSqr, capCoeff, shift, eCap;
};
static_assert(sizeof(Pre) == 80, "table record");
constexpr int MAX_TYPES2 = 64;  // type pairs of a parameter table

// computeForceAndEnergy (LennardJones.hpp:52-78)
__host__ __device__ inline void forceEnergy(const Pre& p, double distSqr, double& ff, double& e)
{
    if (distSqr >= p.capDistSqr)
    {
        const double frac2 = 1.0 / distSqr;
        const double frac6 = frac2 * frac2 * frac2;
        ff = frac6 * (p.ff1 * frac6 - p.ff2) * frac2;
        e = frac6 * (p.ef1 * frac6 - p.ef2) - p.shift;
        return;
    }
    const double dist = sqrt(distSqr);
    ff = p.capCoeff / dist;
    e = p.eCap - (dist - p.capDist) * p.capCoeff - p.shift;
}

// CappedLennardJonesPotential ctor + init functor (LennardJones.cpp:81-116, 61-79)
inline void buildTable(std::vector<Pre>& t, const double* cap, const double* rc, const double* sigma, const double* eps,
                int64_t nt2, bool shifted)
{
    t.assign(size_t(nt2), Pre{});
    for (int64_t k = 0; k < nt2; ++k)
    {
        auto& p = t[size_t(k)];
        const double sig2 = sigma[k] * sigma[k];
        const double sig6 = sig2 * sig2 * sig2;
        p.ff1 = 48.0 * eps[k] * sig6 * sig6;
        p.ff2 = 24.0 * eps[k] * sig6;
        p.ef1 = 4.0 * eps[k] * sig6 * sig6;
        p.ef2 = 4.0 * eps[k] * sig6;
        p.rcSqr = rc[k] * rc[k];
        p.capDist = cap[k];
    }
    for (int64_t k = 0; k < nt2; ++k)
    {
        auto& p = t[size_t(k)];
        const double capDist = p.capDist;
        p.capDist = 0.0;
        p.capDistSqr = 0.0;
        double ff, e;
        forceEnergy(p, capDist * capDist, ff, e);
        p.capCoeff = ff * capDist;
        p.eCap = e;
        p.capDist = capDist;
        p.capDistSqr = capDist * capDist;
        if (shifted)
        {
            forceEnergy(p, p.rcSqr, ff, e);
            p.shift = e;
        }
    }
}

// CheckRegion.hpp:25-38
__host__ __device__ inline bool isAT(double l) { return l >= 1.0; }
__host__ __device__ inline bool isCG(double l) { return l <= 0.0; }
__host__ __device__ inline bool isHY(double l) { return !isAT(l) && !isCG(l); }

// util/math.hpp:31-46
__host__ __device__ inline double powInt(double x, int64_t n)
{
    double ww = x, yy = 1.0;
    for (int64_t nn = (n > 0) ? n : -n; nn != 0; nn >>= 1)
    {
        if ((nn & 1) == 1) yy *= ww;
        ww *= ww;
    }
    return (n > 0) ? yy : 1.0 / yy;
}

// Slab::operator() (weighting_function/Slab.hpp:96-187)
__host__ __device__ inline void slabEval(const mrmd_slab& s, double x, double& lambda, double& mod, double& gx)
{
    const double halfAT = 0.5 * s.atomistic_diameter;
    const double hy = s.hybrid_diameter;
    const int64_t expo = 2 * s.nu;
    const double dx = x - s.center[0];
    const double absDx = fabs(dx);
    if (absDx < halfAT || (!(absDx > halfAT + hy) && s.interface_type == 1))
    {
        lambda = 1.0;
        mod = 1.0;
        gx = 0.0;
        return;
    }
    if (absDx > halfAT + hy)
    {
        lambda = 0.0;
        mod = 0.0;
        gx = 0.0;
        return;
    }
    const double arg = PI / (2.0 * hy) * (absDx - halfAT);
    const double base = cos(arg);
    lambda = base * base;
    mod = powInt(base, expo);
    const double factor = -PI / (2.0 * hy) * double(expo) * sin(arg) * powInt(base, expo - 1) / absDx;
    gx = factor * dx;
}

// weighting_function::Spherical::operator() (Spherical.hpp:37-76): lambda = cos^n of the
// radial position inside the hybrid shell, its gradient along the radius
__host__ __device__ inline void sphericalEval(const mrmd_spherical& s, double x, double y, double z, double& lambda,
                                              double& gx, double& gy, double& gz)
{
    const double dx[3] = {x - s.center[0], y - s.center[1], z - s.center[2]};
    const double dxSqr = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    const double atSqr = s.atomistic_radius * s.atomistic_radius;
    const double cgR = s.atomistic_radius + s.hybrid_diameter;
    if (dxSqr < atSqr)
    {
        lambda = 1.0;
        gx = gy = gz = 0.0;
        return;
    }
    if (dxSqr > cgR * cgR)
    {
        lambda = 0.0;
        gx = gy = gz = 0.0;
        return;
    }
    const double r = sqrt(dxSqr);
    const double arg = PI / (2.0 * s.hybrid_diameter) * (r - s.atomistic_radius);
    const double base = cos(arg);
    lambda = powInt(base, s.exponent);
    const double factor =
        -PI / (2.0 * s.hybrid_diameter) * double(s.exponent) * sin(arg) * powInt(base, s.exponent - 1) / r;
    gx = factor * dx[0];
    gy = factor * dx[1];
    gz = factor * dx[2];
}

// MultiHistogram::getBin (MultiHistogram.hpp:60-66)
__host__ __device__ inline int64_t getBin(double v, double min, double invBinSize, int64_t numBins)
{
    int64_t bin = int64_t(floor((v - min) * invBinSize));
    if (bin < 0) bin = -1;
    if (bin >= numBins) bin = -1;
    return bin;
}

struct HistD
{
    double min, binSize, invBinSize;
    int64_t numBins, numHist;
};

// MultiHistogram.cpp:163-212 for cell (b, h) of `in` (LDS or global)
__device__ inline double smoothAt(const double* in, const HistD& hd, int64_t b, int64_t hh, double sigma, int64_t delta,
                                  int periodic)
{
    const double invSigma = 1.0 / sigma;
    double norm = 0.0, acc = 0.0;
    int64_t jmin = b - delta, jmax = b + delta;
    if (!periodic)
    {
        jmin = jmin > 0 ? jmin : 0;
        jmax = jmax < hd.numBins - 1 ? jmax : hd.numBins - 1;
    }
    for (int64_t j = jmin; j <= jmax; ++j)
    {
        int64_t mj = j;
        if (periodic)
        {
            if (mj < 0) mj += hd.numBins;
            if (mj >= hd.numBins) mj -= hd.numBins;
        }
        const double t = double(b - j) * hd.binSize * invSigma;
        const double eFunc = exp(-(t * t));
        norm += eFunc;
        acc += in[mj * hd.numHist + hh] * eFunc;
    }
    return acc / norm;
}

// MultiHistogram.cpp:113-161
__device__ inline double gradientAt(const double* in, const HistD& hd, int64_t i, int64_t h, int periodic)
{
    const double invSp = hd.invBinSize, invDSp = 0.5 * hd.invBinSize;
    const int64_t nb = hd.numBins, nh = hd.numHist;
    if (i == 0)
        return periodic ? (in[(i + 1) * nh + h] - in[(nb - 1) * nh + h]) * invDSp
                        : (in[(i + 1) * nh + h] - in[i * nh + h]) * invSp;
    if (i == nb - 1)
        return periodic ? (in[h] - in[(i - 1) * nh + h]) * invDSp : (in[i * nh + h] - in[(i - 1) * nh + h]) * invSp;
    return (in[(i + 1) * nh + h] - in[(i - 1) * nh + h]) * invDSp;
}

}  // namespace mrmd_hip
