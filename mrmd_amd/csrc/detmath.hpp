// Counter-based noise and bit-reproducible log/exp for the Langevin O-step.
//
// The reference draws its noise from Kokkos Random_XorShift1024_Pool(1234) with a
// per-thread state assignment that depends on the back-end
// (action/VelocityVerletLangevinThermostat.hpp:32, 101-116).  Here the three normals of
// atom i at call c are a pure function of (seed, i, c): Philox4x32-10 (Salmon, Moraes,
// Dror, Shaw, SC'11) feeding the Marsaglia polar method.  Only +, -, *, / and sqrt (all
// correctly rounded in binary64) are used, and the library is compiled without FMA
// contraction, so any IEEE machine produces identical bits for identical inputs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrmd_hip
{
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r)
    {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        c1 = uint32_t(p1);
        c3 = uint32_t(p0);
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// uniform in (-1, 1) from 52 bits: (m + 1/2) 2^-51 - 1, every operation exact
__host__ __device__ inline double symUniform(uint32_t a, uint32_t b)
{
    const uint64_t m = (uint64_t(a) << 20) ^ uint64_t(b >> 12);
    return (double(m) + 0.5) * 0x1p-51 - 1.0;
}

__host__ __device__ inline double bitsToDouble(uint64_t b)
{
    union
    {
        uint64_t u;
        double d;
    } c;
    c.u = b;
    return c.d;
}
__host__ __device__ inline uint64_t doubleToBits(double d)
{
    union
    {
        uint64_t u;
        double d;
    } c;
    c.d = d;
    return c.u;
}

// log(x), x in (0, 1]: x = m 2^e, m in [1/sqrt2, sqrt2), log m = 2 atanh(t), t = (m-1)/(m+1)
__host__ __device__ inline double detLog(double x)
{
    const uint64_t bits = doubleToBits(x);
    int e = int((bits >> 52) & 0x7ff) - 1023;
    double m = bitsToDouble((bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    if (m > 1.4142135623730951)
    {
        m = m * 0.5;
        e += 1;
    }
    const double t = (m - 1.0) / (m + 1.0);
    const double t2 = t * t;
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

// exp(x), x <= 0: x = k ln2 + r (two-part ln2), Taylor series of e^r to degree 13, times 2^k
__host__ __device__ inline double detExp(double x)
{
    if (x < -700.0) return 0.0;
    if (x == 0.0) return 1.0;
    const double kd = floor(x * 1.4426950408889634 + 0.5);
    const double r = (x - kd * 6.93147180369123816490e-01) - kd * 1.90821492927058770002e-10;
    double p = 1.0;
    for (int n = 13; n >= 1; --n) p = 1.0 + (p * r) / double(n);
    const int k = int(kd);
    return p * bitsToDouble(uint64_t(k + 1023) << 52);
}

// three standard normals for (seed, id, call); replica mixes into the key's high word
__host__ __device__ inline void normals3(uint64_t seed, uint64_t id, uint64_t call, double out[3])
{
    const uint32_t k0 = uint32_t(seed), k1 = uint32_t(seed >> 32);
    int have = 0;
    for (uint32_t round = 0; have < 3; ++round)
    {
        uint32_t r[4];
        philox4x32_10(uint32_t(id), uint32_t(call), uint32_t(call >> 32), round, k0, k1, r);
        const double u = symUniform(r[0], r[1]);
        const double v = symUniform(r[2], r[3]);
        const double s = u * u + v * v;
        if (!(s < 1.0 && s > 0.0)) continue;
        const double f = sqrt(-2.0 * detLog(s) / s);
        out[have++] = u * f;
        if (have < 3) out[have++] = v * f;
    }
}

// The normals of normals3 for every lane of a 64-lane wave, with the same Philox draws, the
// same polar acceptances in the same round order and the same arithmetic, hence the same
// bits, but the work regrouped for the wave:
//  1. every lane draws rounds 0-2 (88 % of lanes hold two accepted pairs after them);
//  2. lanes still short of two pairs are served by the whole wave: each lane draws one
//     (short lane, round) item of the following rounds, the draws pass through LDS, and
//     each short lane takes its accepted pairs in round order, until none is short;
//  3. the logarithm, division and square root of the transform run once per pair with the
//     wave converged, not in every pass of a divergent rejection loop.
// Every lane of the wave calls it with its own (seed, id) (need = false where no normals
// are wanted); `call` is the same for the whole wave.
struct WaveDraws
{
    uint64_t seed[64], id[64];
    double u[64], v[64];
};

// beyond this round a lane keeps zero normals; it is not reached (a lane is short after
// r rounds with probability below 0.22^(r-1))
constexpr uint32_t MAX_NORMAL_ROUNDS = 4096;

__device__ inline void waveNormals3(uint64_t seed, uint64_t id, uint64_t call, bool need, WaveDraws& w,
                                    double out[3])
{
    const uint32_t c1 = uint32_t(call), c2 = uint32_t(call >> 32);
    const int lane = int(__lane_id());
    double u0 = 0.0, v0 = 0.0, s0 = 1.0, u1 = 0.0, v1 = 0.0, s1 = 1.0;
    int have = need ? 0 : 2;
    auto take = [&](double u, double v, double s) {
        if (have == 0)
        {
            u0 = u;
            v0 = v;
            s0 = s;
        }
        else
        {
            u1 = u;
            v1 = v;
            s1 = s;
        }
        ++have;
    };
#pragma unroll
    for (uint32_t r = 0; r < 3; ++r)
    {
        uint32_t o[4];
        philox4x32_10(uint32_t(id), c1, c2, r, uint32_t(seed), uint32_t(seed >> 32), o);
        const double u = symUniform(o[0], o[1]);
        const double v = symUniform(o[2], o[3]);
        const double s = u * u + v * v;
        if (s < 1.0 && s > 0.0 && have < 2) take(u, v, s);
    }
    // every short lane has been served the same rounds so far: `next` is wave-uniform
    for (uint32_t next = 3; next < MAX_NORMAL_ROUNDS;)
    {
        const bool shortOf = have < 2;
        const uint64_t m = __ballot(shortOf);
        if (m == 0) break;
        const int ns = __popcll(m);
        const int rank = __popcll(m & ((uint64_t(1) << lane) - 1));
        if (shortOf)
        {
            w.seed[rank] = seed;
            w.id[rank] = id;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int per = 64 / ns;  // rounds drawn for each short lane in this pass
        const int k = lane / per, e = lane % per;
        bool ok = false;
        if (k < ns)
        {
            const uint64_t sk = w.seed[k];
            uint32_t o[4];
            philox4x32_10(uint32_t(w.id[k]), c1, c2, next + uint32_t(e), uint32_t(sk), uint32_t(sk >> 32), o);
            const double u = symUniform(o[0], o[1]);
            const double v = symUniform(o[2], o[3]);
            const double s = u * u + v * v;
            ok = s < 1.0 && s > 0.0;
            w.u[lane] = u;
            w.v[lane] = v;
        }
        const uint64_t okm = __ballot(ok);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (shortOf)
            for (int j = 0; j < per && have < 2; ++j)
            {
                const int src = rank * per + j;
                if ((okm >> src) & 1)
                {
                    const double u = w.u[src], v = w.v[src];
                    take(u, v, u * u + v * v);
                }
            }
        next += uint32_t(per);
        // the next pass rewrites w
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    const double f0 = sqrt(-2.0 * detLog(s0) / s0);
    const double f1 = sqrt(-2.0 * detLog(s1) / s1);
    out[0] = u0 * f0;
    out[1] = v0 * f0;
    out[2] = u1 * f1;
}
}  // namespace mrmd_hip
