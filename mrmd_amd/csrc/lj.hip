// Lennard-Jones pair force (action::LennardJones), its precomputed table, and the adjoint.
//
// Force kernel design (gfx950, FP64): one lane per local atom over its full neighbour row
// (both directions of every pair), so F_i is owned by one lane and no atomics are needed;
// the pair work is done twice, which on a chip with 78.6 TFLOP/s FP64 costs less than the
// L2 atomics of a half list.  One workgroup per Verlet-list tile stages the neighbourhood's
// {x,y,z,type} records in LDS once; rows then read neighbours from LDS through 16-bit slot
// entries.  Energy and virial are reduced with wave shuffles, per-workgroup partials and
// one ordered final pass (deterministic).
// The reference's half-list algorithm (atomic F_j, LennardJones.hpp:158-205) is kept for
// asymmetric pair predicates and asymmetric parameter tables.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "nlist.hpp"
#include "physics.hpp"

using namespace mrmd_hip;

namespace
{
}  // namespace

struct mrmd_lj
{
    mrmd_hip::CtxPtr ctx;
    int64_t numTypes = 1, typeStride = 1;
    bool shifted = false;
    bool symmetric = true;  ///< table(ti*stride+tj) == table(tj*stride+ti) for all types
    std::vector<double> cap, rc;
    std::vector<double> sigma, eps;  ///< the parameters the table was built from
    std::vector<Pre> table;
    double rcSqr = 0.0;
    DBuf<Pre> dtable;
    DBuf<double> dParams;  ///< sigma then epsilon per table entry (the adjoint's parameters)
    DBuf<double> ev;  ///< energy, virial on the device
    /// adjoint work buffers, kept between calls: upstream gradient (component-major over all
    /// records), dL/dx (component-major), per-workgroup parameter partials, row-major staging
    DBuf<double> adjG, adjOut, adjPart, adjStage;
    DBuf<double> partials;   ///< per-workgroup (energy, virial) of the last apply
    int64_t nPartials = 0;   ///< > 0: partials not yet summed into ev
    bool evValid = true;
    double hE = 0.0, hW = 0.0;

    void upload();
};

// true when mrmd_lj_apply without a predicate runs the full-row tiled kernel, which writes
// the forces of list rows only (ghost records receive none)
bool mrmd_hip::ljFullRows(const mrmd_lj* lj) { return lj && lj->symmetric; }

void mrmd_lj::upload()
{
    dtable.reserve(ctx, int64_t(table.size()));
    HIPCHECK(hipMemcpyAsync(dtable.p, table.data(), table.size() * sizeof(Pre), hipMemcpyHostToDevice, ctx->stream));
    const size_t nt2 = table.size();
    dParams.reserve(ctx, int64_t(2 * nt2));
    HIPCHECK(hipMemcpyAsync(dParams.p, sigma.data(), nt2 * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipMemcpyAsync(dParams.p + nt2, eps.data(), nt2 * 8, hipMemcpyHostToDevice, ctx->stream));
    sync(ctx);
    symmetric = true;
    for (int64_t a = 0; a < numTypes; ++a)
        for (int64_t b = 0; b < numTypes; ++b)
        {
            const int64_t i1 = a * typeStride + b, i2 = b * typeStride + a;
            if (i1 >= int64_t(table.size()) || i2 >= int64_t(table.size())) continue;
            if (std::memcmp(&table[size_t(i1)], &table[size_t(i2)], sizeof(Pre)) != 0) symmetric = false;
        }
}

namespace
{
struct LJArgs
{
    const double4* pos;
    double* force;
    int64_t cap;
    int64_t begin;
    int64_t rows;
    int64_t stride;
    const Pre* table;
    int nt2;
    int typeStride;
    double rcSqr;
    mrmd_pred2 pred;
    int usePred;
};

// Tiled full-row force: one workgroup per list tile stages the current positions of the
// tile's neighbourhood (sorted-slot order) in LDS; each row reads its neighbours from LDS
// through the 16-bit slot entries, 8 entries per 16-byte load with the next chunk's load
// issued before the current chunk is processed.  ONE: single-type table held in registers.
struct TiledList
{
    const uint16_t* nbr;
    const int32_t* cnt;
    int64_t strideS;
    const TileInfo* tiles;
    const int32_t* sorted;
    const double4* sortedPos;
    int32_t maxStaged;  ///< the sentinel record sits at sentinelSlot(maxStaged)
    double sentinel;    ///< at (sentinel, sentinel, sentinel)
    double4* gstage;    ///< the list's per-tile global staging (dense inputs), or null: LDS
    int64_t gstageRecs; ///< records per tile slice of gstage
};

// 1/d^2 outside the cutoff made negligible with one select: the high word cleared leaves a
// value below 2^-1022 whose square underflows to zero, so every term of the pair (f2^3, its
// products and the force) comes out zero as from a zeroed reciprocal, with the same signs
__device__ inline double maskReciprocal(double f2, bool in)
{
    const long long b = __double_as_longlong(f2);
    return __hiloint2double(in ? int(b >> 32) : 0, int(b));
}

// Branch-free pair terms of one chunk of eight entries, written stage by stage over the
// eight pairs so that the eight dependency chains (LDS read, distance, reciprocal, force)
// are independent and interleave.  Padding entries (the sentinel record) and pairs beyond
// rc get a zero factor.  CAP: the table has a capping distance, so the capped branch
// (LennardJones.hpp:68-75) is evaluated too (single-type tables: only for the chunks where
// a lane of the wave needs it).  Products and sums may contract to FMAs here
// (the neighbour decision of the list is unaffected: it is made, unfused, at build time);
// the forces then differ from the unfused CPU arithmetic by a few ulps per term.
template <bool PRED, bool ONE, bool CAP>
__device__ inline void ljChunk(const LJArgs& a, const Pre* tab, const Pre& p1, int ti0, const double4& pi,
                               const void* st, const uint4 cur, double& fx, double& fy, double& fz, double& e,
                               double& w, int& n)
{
#pragma clang fp contract(fast)
    const uint32_t wd[4] = {cur.x, cur.y, cur.z, cur.w};
    double dx[CHUNK], dy[CHUNK], dz[CHUNK], d2[CHUNK], ty[CHUNK];
#pragma unroll
    for (int u = 0; u < CHUNK; ++u)
    {
        // single-type tables stage {x, y, z} only
        const double4 pj = Staged<ONE>::get(st, int((wd[u >> 1] >> ((u & 1) * 16)) & 0x7fffu));
        dx[u] = pi.x - pj.x;
        dy[u] = pi.y - pj.y;
        dz[u] = pi.z - pj.z;
        ty[u] = pj.w;
        if (PRED && !evalPred2(a.pred, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z)) ty[u] = -1.0;
    }
#pragma unroll
    for (int u = 0; u < CHUNK; ++u) d2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
    double frac2[CHUNK];
#pragma unroll
    for (int u = 0; u < CHUNK; ++u) frac2[u] = recipNewton(d2[u]);
    auto accumulate = [&](int u, double ff, double en) {
        // padding entries point at the sentinel record beyond every cutoff, so the cutoff test
        // alone decides; the terms are finite for every entry and a 0 / 1 factor keeps them
        bool in = d2[u] <= a.rcSqr;
        if (PRED) in = in && ty[u] >= 0.0;
        const double m = in ? 1.0 : 0.0;
        ff *= m;
        en *= m;
        fx += dx[u] * ff;
        fy += dy[u] * ff;
        fz += dz[u] * ff;
        e += en;
        w += ff * d2[u];
    };
    if constexpr (ONE && !CAP)
    {
        // single-type table without capping: the reciprocal is masked (0 outside rc), and
        // instead of each pair's energy and virial the lane keeps e = sum f6, w = sum g with
        // g = f6 (ff1 f6 - ff2) = ff d^2, and n = pairs inside rc; the kernel turns them into
        // energy ef1 sum f6^2 - ef2 sum f6 - shift n, sum f6^2 = (sum g + ff2 sum f6) / ff1
#pragma unroll
        for (int u = 0; u < CHUNK; ++u)
        {
            bool in = d2[u] <= a.rcSqr;
            if (PRED) in = in && ty[u] >= 0.0;
            const double f2 = maskReciprocal(frac2[u], in);
            const double f6 = f2 * f2 * f2;
            const double g = f6 * (p1.ff1 * f6 - p1.ff2);
            const double ff = g * f2;
            fx += dx[u] * ff;
            fy += dy[u] * ff;
            fz += dz[u] * ff;
            e += f6;
            w += g;
            unsigned carry;  // the count takes the comparison as a carry (one add)
            n = int(__builtin_addc(unsigned(n), 0u, unsigned(in), &carry));
        }
        return;
    }
    if constexpr (!CAP || !ONE)
    {
        // multi-type tables keep the capped branch per pair (the split form below would hold
        // two table records per pair and cost occupancy)
#pragma unroll
        for (int u = 0; u < CHUNK; ++u)
        {
            const Pre& p = ONE ? p1 : tab[ti0 + max(int(ty[u]), 0)];
            const double frac6 = frac2[u] * frac2[u] * frac2[u];
            double ff = frac6 * (p.ff1 * frac6 - p.ff2) * frac2[u];
            double en = frac6 * (p.ef1 * frac6 - p.ef2) - p.shift;
            if (CAP && d2[u] < p.capDistSqr)
            {
                const double dist = sqrt(d2[u]);
                ff = p.capCoeff / dist;
                en = p.eCap - (dist - p.capDist) * p.capCoeff - p.shift;
            }
            accumulate(u, ff, en);
        }
    }
    else
    {
        double ffv[CHUNK], env[CHUNK];
        bool capped = false;
#pragma unroll
        for (int u = 0; u < CHUNK; ++u)
        {
            const Pre& p = ONE ? p1 : tab[ti0 + max(int(ty[u]), 0)];
            const double frac6 = frac2[u] * frac2[u] * frac2[u];
            ffv[u] = frac6 * (p.ff1 * frac6 - p.ff2) * frac2[u];
            env[u] = frac6 * (p.ef1 * frac6 - p.ef2) - p.shift;
            capped = capped || d2[u] < p.capDistSqr;
        }
        // pairs inside the capping distance are rare: the capped form runs, by selects over
        // the chunk, only where some lane of the wave has one
        if (__any(capped))
        {
            // keeps the block a branch: the compiler would otherwise evaluate it every chunk
            asm volatile("" ::: "memory");
#pragma unroll
            for (int u = 0; u < CHUNK; ++u)
            {
                const Pre& p = ONE ? p1 : tab[ti0 + max(int(ty[u]), 0)];
                const double dist = sqrt(d2[u]);
                const bool c = d2[u] < p.capDistSqr;
                ffv[u] = c ? p.capCoeff / dist : ffv[u];
                env[u] = c ? p.eCap - (dist - p.capDist) * p.capCoeff - p.shift : env[u];
            }
        }
#pragma unroll
        for (int u = 0; u < CHUNK; ++u) accumulate(u, ffv[u], env[u]);
    }
}

// ASSIGN: the forces of the rows are stored as 0.0 + F instead of added (the caller zeroed
// nothing there; 0.0 + F is what zeroing and adding leave, -0.0 included)
// GS: the neighbourhood is staged in the tile's slice of the list's global staging buffer
// (dense inputs whose tiles exceed the LDS) instead of LDS
template <bool PRED, bool ONE, bool CAP, bool ASSIGN = false, bool GS = false>
__global__ __launch_bounds__(BLOCK) void k_ljTiled(LJArgs a, TiledList tl, Pre p1, double* partials)
{
    extern __shared__ double4 stLds[];
    __shared__ Pre tab[ONE ? 1 : MAX_TYPES2];
    __shared__ TileInfo ti;
    __shared__ double red[BLOCK / WAVE];
    const unsigned blk = xcdBlock();
    double4* const st = GS ? tl.gstage + int64_t(blk) * tl.gstageRecs : stLds;
    if (!ONE)
        for (int t = threadIdx.x; t < a.nt2; t += blockDim.x) tab[t] = a.table[t];
    if (threadIdx.x == 0) ti = tl.tiles[blk];
    __syncthreads();
    // the first row's global loads (count, index, position, first index chunk) are issued
    // before the staging so that their latency overlaps it
    int s = -1, c = 0;
    int32_t i = 0;
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
    uint4 cur = {0, 0, 0, 0};
    if (int(threadIdx.x) < ti.nrows)
    {
        int self;
        s = tileRow(ti, threadIdx.x, self);
        c = tl.cnt[s];
        i = tl.sorted[s];
        pi = tl.sortedPos[s];
        cur = reinterpret_cast<const uint4*>(tl.nbr)[s];
    }
    stageTile<ONE>(ti, tl.sortedPos, st);
    putSentinel<ONE>(st, tl.maxStaged, tl.sentinel, 0.0);
    __syncthreads();
    double e = 0.0, w = 0.0;
    int n = 0;
    for (int t = threadIdx.x; t < ti.nrows; t += blockDim.x)
    {
        if (t != int(threadIdx.x))
        {
            int self;
            s = tileRow(ti, t, self);
            c = tl.cnt[s];
            i = tl.sorted[s];
            pi = tl.sortedPos[s];
            cur = reinterpret_cast<const uint4*>(tl.nbr)[s];
        }
        if ((!ASSIGN && c == 0) || i >= a.rows) continue;
        const int ti0 = ONE ? 0 : int(pi.w) * a.typeStride;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        const uint4* chunk = reinterpret_cast<const uint4*>(tl.nbr) + s;
        const int nch = (c + CHUNK - 1) / CHUNK;
        for (int ch = 0; ch < nch; ++ch)
        {
            uint4 nxt = cur;
            if (ch + 1 < nch) nxt = chunk[int64_t(ch + 1) * tl.strideS];
            ljChunk<PRED, ONE, CAP>(a, tab, p1, ti0, pi, st, cur, fx, fy, fz, e, w, n);
            cur = nxt;
        }
        if (ASSIGN)
        {
            a.force[i] = 0.0 + fx;
            a.force[a.cap + i] = 0.0 + fy;
            a.force[2 * a.cap + i] = 0.0 + fz;
        }
        else
        {
            a.force[i] += fx;
            a.force[a.cap + i] += fy;
            a.force[2 * a.cap + i] += fz;
        }
    }
    if constexpr (ONE && !CAP)
    {
        // e = sum f6, w = sum g over this lane's pairs (see ljChunk)
        const double s12 = p1.ff1 != 0.0 ? (w + p1.ff2 * e) / p1.ff1 : 0.0;
        e = p1.ef1 * s12 - p1.ef2 * e - p1.shift * double(n);
    }
    // every pair is visited from both ends: E = 1/2 sum, W = -1/2 sum(ff d2) / 2
    e = blockSum(0.5 * e, red);
    w = blockSum(-0.25 * w, red);
    if (threadIdx.x == 0)
    {
        partials[2 * blk] = e;
        partials[2 * blk + 1] = w;
    }
}

// the reference algorithm: half rows, F_j by atomics (LennardJones.hpp:158-205)
template <bool PRED>
__global__ __launch_bounds__(BLOCK) void k_ljHalf(LJArgs a, const int32_t* __restrict__ nbr,
                                                  const int32_t* __restrict__ cnt, double* partials)
{
    __shared__ Pre tab[MAX_TYPES2];
    __shared__ double red[BLOCK / WAVE];
    for (int t = threadIdx.x; t < a.nt2; t += blockDim.x) tab[t] = a.table[t];
    __syncthreads();
    const unsigned blk = xcdBlock();
    const int64_t r = int64_t(blk) * blockDim.x + threadIdx.x;
    double e = 0.0, w = 0.0;
    if (r < a.rows)
    {
        const int64_t i = a.begin + r;
        const double4 pi = a.pos[i];
        const int ti = int(pi.w) * a.typeStride;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        const int c = cnt[r];
        for (int k = 0; k < c; ++k)
        {
            const int32_t j = nbr[int64_t(k) * a.stride + r];
            const double4 pj = a.pos[j];
            if (PRED && !evalPred2(a.pred, pi.x, pi.y, pi.z, pj.x, pj.y, pj.z)) continue;
            const double dx = pi.x - pj.x;
            const double dy = pi.y - pj.y;
            const double dz = pi.z - pj.z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > a.rcSqr) continue;
            const Pre& p = tab[ti + int(pj.w)];
            double ff, en;
            forceEnergy(p, d2, ff, en);
            e += en;
            w -= 0.5 * ff * d2;
            fx += dx * ff;
            fy += dy * ff;
            fz += dz * ff;
            atomicAddD(&a.force[j], -(dx * ff));
            atomicAddD(&a.force[a.cap + j], -(dy * ff));
            atomicAddD(&a.force[2 * a.cap + j], -(dz * ff));
        }
        atomicAddD(&a.force[i], fx);
        atomicAddD(&a.force[a.cap + i], fy);
        atomicAddD(&a.force[2 * a.cap + i], fz);
    }
    e = blockSum(e, red);
    w = blockSum(w, red);
    if (threadIdx.x == 0)
    {
        partials[2 * blk] = e;
        partials[2 * blk + 1] = w;
    }
}

// ordered sum of per-workgroup partials (one workgroup)
__global__ void k_sumPartials(const double* partials, int64_t n, int width, double* out)
{
    __shared__ double red[BLOCK / WAVE];
    for (int c = 0; c < width; ++c)
    {
        double s = 0.0;
        for (int64_t k = threadIdx.x; k < n; k += blockDim.x) s += partials[k * width + c];
        s = blockSum(s, red);
        if (threadIdx.x == 0) out[c] = s;
    }
}

__global__ void k_ljEval(const Pre* table, const double* d2, const int64_t* idx, int64_t n, double* ff, double* e)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    forceEnergy(table[idx[k]], d2[k], ff[k], e[k]);
}

// ---------------------------------------------------------------------------------------
// Adjoint of F_i = sum_j s(r_ij^2) d_ij, d = x_i - x_j, s = r^-8 (ff1 r^-6 - ff2):
//   dL/dx_i = sum_j J_ij (g_i - g_j),  J = s I + 2 s' d d^T,  s' = ds/d(r^2)
//   dL/deps_ab   = sum_{pairs ab} (g_i - g_j) . d s/eps
//   dL/dsigma_ab = sum_{pairs ab} (g_i - g_j) . d ds/dsigma
// With full rows every pair is seen from both ends: dL/dx_i is complete per row, and the
// parameter sums take half of each row's term.  Capped pairs use the capped expression.
constexpr int ADJ_BLOCK = 64;  // small workgroups: a 32k-atom snapshot still fills every CU
constexpr int ADJ_LANES = 4;   // lanes per row: each takes every 4th neighbour, shuffles combine
constexpr int ADJ_CHUNK = 2;   // neighbours a lane evaluates together (independent streams)

template <bool FEW, bool CAP>
__global__ __launch_bounds__(ADJ_BLOCK) void k_ljAdjoint(LJArgs a, const uint32_t* __restrict__ nbr,
                                                         const int32_t* __restrict__ cnt,
                                                         const double* __restrict__ g, int64_t gstride,
                                                         const double* __restrict__ sigma,
                                                         const double* __restrict__ eps, double* dLdx,
                                                         int64_t xstride, double* pgrad /*[block][2*nt2]*/)
{
    __shared__ Pre tab[MAX_TYPES2];
    __shared__ double isg[MAX_TYPES2], ise[MAX_TYPES2];
    __shared__ double acc[2 * MAX_TYPES2];
    for (int t = threadIdx.x; t < a.nt2; t += blockDim.x)
    {
        tab[t] = a.table[t];
        isg[t] = 1.0 / sigma[t];
        ise[t] = 1.0 / eps[t];
        acc[t] = 0.0;
        acc[MAX_TYPES2 + t] = 0.0;
    }
    __syncthreads();
    // FEW: up to 4 type pairs accumulate in registers, otherwise in LDS with atomics
    double re[4] = {0.0, 0.0, 0.0, 0.0}, rs[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t gt = int64_t(xcdBlock()) * blockDim.x + threadIdx.x;
    const int64_t r = gt / ADJ_LANES;
    const int sub = int(gt % ADJ_LANES);
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if (r < a.rows)
    {
        const int64_t i = a.begin + r;
        const double4 pi = a.pos[i];
        const int ti = int(pi.w) * a.typeStride;
        const double gix = g[r], giy = g[gstride + r], giz = g[2 * gstride + r];
        const int c = cnt[r];
        for (int k0 = sub; k0 < c; k0 += ADJ_LANES * ADJ_CHUNK)
        {
            // entries past the row end read the row itself: d^2 = 0 selects zero below
            int64_t jr[ADJ_CHUNK];
            double4 pj[ADJ_CHUNK];
#pragma unroll
            for (int u = 0; u < ADJ_CHUNK; ++u)
            {
                const int k = k0 + u * ADJ_LANES;
                const int64_t j = k < c ? int64_t(nbr[int64_t(k) * a.stride + r] & 0x7fffffffu) : i;
                jr[u] = j - a.begin;
                pj[u] = a.pos[j];
            }
#pragma unroll
            for (int u = 0; u < ADJ_CHUNK; ++u)
            {
                const double dx = pi.x - pj[u].x;
                const double dy = pi.y - pj[u].y;
                const double dz = pi.z - pj[u].z;
                const double d2 = dx * dx + dy * dy + dz * dz;
                const bool in = d2 <= a.rcSqr && d2 > 0.0;
                const int tIdx = ti + int(pj[u].w);
                const Pre& p = tab[tIdx];
                // upstream of the partner: ghosts carry their real atom's upstream (expanded g)
                const double ux = gix - g[jr[u]], uy = giy - g[gstride + jr[u]], uz = giz - g[2 * gstride + jr[u]];
                const double du = dx * ux + dy * uy + dz * uz;
                const double f2 = 1.0 / d2;
                const double f6 = f2 * f2 * f2;
                double sv = f6 * (p.ff1 * f6 - p.ff2) * f2;
                // ds/d(r^2) = -(7 ff1 r^-16 - 4 ff2 r^-10)
                double sp = -(7.0 * p.ff1 * f6 * f6 * f2 * f2 - 4.0 * p.ff2 * f6 * f2 * f2);
                double dsds = f6 * (12.0 * p.ff1 * f6 - 6.0 * p.ff2) * f2 * isg[tIdx];
                if (CAP && d2 < p.capDistSqr)
                {
                    const double dist = sqrt(d2);
                    sv = p.capCoeff / dist;
                    sp = -0.5 * p.capCoeff / (dist * d2);
                    // capCoeff = s(rcap) rcap is linear in eps; its sigma-derivative follows s
                    const double c2 = 1.0 / p.capDistSqr;
                    const double c6 = c2 * c2 * c2;
                    dsds = (c6 * (12.0 * p.ff1 * c6 - 6.0 * p.ff2) * c2 * isg[tIdx]) * p.capDist / dist;
                }
                sv = in ? sv : 0.0;
                sp = in ? sp : 0.0;
                dsds = in ? dsds : 0.0;
                const double dsde = sv * ise[tIdx];
                const double two_sp_du = 2.0 * sp * du;
                ox += sv * ux + two_sp_du * dx;
                oy += sv * uy + two_sp_du * dy;
                oz += sv * uz + two_sp_du * dz;
                const double h = 0.5 * du;
                if (FEW)
                {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                    {
                        re[q] += q == tIdx ? h * dsde : 0.0;
                        rs[q] += q == tIdx ? h * dsds : 0.0;
                    }
                }
                else if (in)
                {
                    atomicAdd(&acc[tIdx], h * dsde);
                    atomicAdd(&acc[MAX_TYPES2 + tIdx], h * dsds);
                }
            }
        }
    }
    // the row's lanes are adjacent in the wave: combine their partial sums
#pragma unroll
    for (int off = 1; off < ADJ_LANES; off <<= 1)
    {
        ox += __shfl_xor(ox, off, WAVE);
        oy += __shfl_xor(oy, off, WAVE);
        oz += __shfl_xor(oz, off, WAVE);
    }
    if (r < a.rows && sub == 0)
    {
        dLdx[r] = ox;
        dLdx[xstride + r] = oy;
        dLdx[2 * xstride + r] = oz;
    }
    // per-workgroup partial parameter gradients, summed in order by k_sumPartials
    double* part = pgrad + int64_t(blockIdx.x) * 2 * a.nt2;
    if (FEW)
    {
        __shared__ double red[ADJ_BLOCK / WAVE];
        for (int q = 0; q < a.nt2; ++q)
        {
            const double se_ = blockSum(re[q], red);
            const double ss_ = blockSum(rs[q], red);
            if (threadIdx.x == 0)
            {
                part[q] = se_;
                part[a.nt2 + q] = ss_;
            }
        }
        return;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < a.nt2; t += blockDim.x)
    {
        part[t] = acc[t];
        part[a.nt2 + t] = acc[MAX_TYPES2 + t];
    }
}

// upstream gradient of ghosts = upstream of their real atom: expand g (local rows) to all
// records [begin, nall) so the adjoint can read g_j for ghost partners
__global__ void k_expandGhostUpstream(double* g, int64_t stride, const int32_t* corr, int64_t nlocal, int64_t nall)
{
    const int64_t i = nlocal + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nall) return;
    const int32_t rr = corr[i];
    for (int d = 0; d < 3; ++d) g[d * stride + i] = g[d * stride + rr];
}

__global__ void k_zeroD(double* p, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0;
}

__global__ void k_rowMajorToSoA(const double* in, double* out, int64_t n, int64_t stride)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    for (int d = 0; d < 3; ++d) out[d * stride + k] = in[3 * k + d];
}
__global__ void k_soaToRowMajor(const double* in, double* out, int64_t n, int64_t stride)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    for (int d = 0; d < 3; ++d) out[3 * k + d] = in[d * stride + k];
}
}  // namespace


extern "C" {

int mrmd_lj_create(mrmd_ctx* c, const double* cap, const double* rc, const double* sigma, const double* eps,
                   int64_t numTypes, int shifted, int64_t typeStride, mrmd_lj** out)
{
    MRMD_API_BEGIN(c)
    REQUIRE(out && cap && rc && sigma && eps && numTypes >= 1, "invalid arguments");
    REQUIRE(numTypes * numTypes <= MAX_TYPES2, "at most %d type pairs", MAX_TYPES2);
    REQUIRE(typeStride >= 1, "type stride must be positive");
    auto* lj = new mrmd_lj();
    lj->ctx = c;
    lj->numTypes = numTypes;
    lj->typeStride = typeStride;
    lj->shifted = shifted != 0;
    const int64_t nt2 = numTypes * numTypes;
    lj->cap.assign(cap, cap + nt2);
    lj->rc.assign(rc, rc + nt2);
    lj->sigma.assign(sigma, sigma + nt2);
    lj->eps.assign(eps, eps + nt2);
    buildTable(lj->table, cap, rc, sigma, eps, nt2, shifted != 0);
    double rcMax = rc[0];
    for (int64_t k = 0; k < nt2; ++k) rcMax = std::max(rcMax, rc[k]);
    lj->rcSqr = rcMax * rcMax;
    try
    {
        lj->upload();
        lj->ev.reserve(c, 2);
        hipLaunchKernelGGL(k_zeroD, dim3(1), dim3(64), 0, c->stream, lj->ev.p, int64_t(2));
        checkLaunch();
    }
    catch (...)
    {
        delete lj;
        throw;
    }
    *out = lj;
    MRMD_API_END
}

int mrmd_lj_destroy(mrmd_lj* lj)
{
    if (!lj) return MRMD_OK;
    (void)hipStreamSynchronize(lj->ctx->stream);
    delete lj;
    return MRMD_OK;
}

int mrmd_lj_set_params(mrmd_lj* lj, const double* sigma, const double* eps)
{
    MRMD_API_BEGIN(lj ? lj->ctx : nullptr)
    REQUIRE(lj && sigma && eps, "invalid arguments");
    const int64_t nt2 = int64_t(lj->table.size());
    buildTable(lj->table, lj->cap.data(), lj->rc.data(), sigma, eps, nt2, lj->shifted);
    lj->sigma.assign(sigma, sigma + nt2);
    lj->eps.assign(eps, eps + nt2);
    lj->upload();
    MRMD_API_END
}

int mrmd_lj_table(const mrmd_lj* lj, double* host)
{
    if (!lj || !host) return MRMD_ERR_ARG;
    std::memcpy(host, lj->table.data(), lj->table.size() * sizeof(Pre));
    return MRMD_OK;
}

int mrmd_lj_eval(mrmd_lj* lj, const double* d2, const int64_t* idx, int64_t n, double* ff, double* e)
{
    MRMD_API_BEGIN(lj ? lj->ctx : nullptr)
    REQUIRE(lj && d2 && idx && ff && e && n >= 0, "invalid arguments");
    for (int64_t k = 0; k < n; ++k)
        REQUIRE(idx[k] >= 0 && idx[k] < int64_t(lj->table.size()), "type index %lld out of range", (long long)idx[k]);
    if (n == 0) return MRMD_OK;
    mrmd_ctx* ctx = lj->ctx;
    double *dd, *dff, *de;
    int64_t* di;
    HIPCHECK(hipMalloc(&dd, size_t(n) * 8));
    HIPCHECK(hipMalloc(&dff, size_t(n) * 8));
    HIPCHECK(hipMalloc(&de, size_t(n) * 8));
    HIPCHECK(hipMalloc(&di, size_t(n) * 8));
    HIPCHECK(hipMemcpyAsync(dd, d2, size_t(n) * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPCHECK(hipMemcpyAsync(di, idx, size_t(n) * 8, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_ljEval, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, lj->dtable.p, dd, di, n, dff, de);
    checkLaunch();
    HIPCHECK(hipMemcpyAsync(ff, dff, size_t(n) * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHECK(hipMemcpyAsync(e, de, size_t(n) * 8, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    (void)hipFree(dd);
    (void)hipFree(dff);
    (void)hipFree(de);
    (void)hipFree(di);
    MRMD_API_END
}

}  // extern "C"

namespace
{
// the body of mrmd_lj_apply; assign: store the rows' forces instead of adding them, honoured
// only by the full-row kernel of a single-type table without capping and predicate
// (returns whether it was)
bool ljApplyImpl(mrmd_lj* lj, mrmd_atoms* at, mrmd_verlet_list* vl, const mrmd_pred2* pred, bool assign);
}  // namespace

bool mrmd_hip::ljAssignable(const mrmd_lj* lj)
{
    if (!lj || !lj->symmetric || lj->table.size() != 1) return false;
    return !(lj->table[0].capDistSqr > 0.0);
}

int mrmd_hip::ljApplyAssign(mrmd_lj* lj, mrmd_atoms* at, mrmd_verlet_list* vl)
{
    MRMD_API_BEGIN(lj ? lj->ctx : nullptr)
    REQUIRE(ljAssignable(lj), "the table has no store form");
    // without rows nothing is stored (and nothing needs to be)
    if (!ljApplyImpl(lj, at, vl, nullptr, true) && at && vl && std::min<int64_t>(at->nlocal, vl->end) > 0)
        fail(MRMD_ERR_STATE, "the store form did not run");
    MRMD_API_END
}

extern "C" int mrmd_lj_apply(mrmd_lj* lj, mrmd_atoms* at, mrmd_verlet_list* vl, const mrmd_pred2* pred)
{
    MRMD_API_BEGIN(lj ? lj->ctx : nullptr)
    ljApplyImpl(lj, at, vl, pred, false);
    MRMD_API_END
}

namespace
{
bool ljApplyImpl(mrmd_lj* lj, mrmd_atoms* at, mrmd_verlet_list* vl, const mrmd_pred2* pred, bool assign)
{
    bool assigned = false;
    REQUIRE(lj && at && vl && at->ctx == lj->ctx && vl->ctx == lj->ctx, "invalid arguments");
    validatePred2(pred);
    REQUIRE(vl->begin == 0 && vl->end <= at->nlocal + at->nghost, "the list rows must start at record 0");
    REQUIRE(vl->nrows <= at->size, "the list was built for more records than the atoms hold");
    REQUIRE(at->maxType * lj->typeStride + at->maxType < int64_t(lj->table.size()),
            "atom type %lld has no entry in the %lld-type table", (long long)at->maxType, (long long)lj->numTypes);
    // the reference loops over local atoms with the list rows (LennardJones.hpp:146)
    const int64_t rows = std::min<int64_t>(at->nlocal, vl->end);
    const bool fresh = vl->sortedFresh;  // consumed by this call whatever path it takes
    vl->sortedFresh = false;
    mrmd_ctx* ctx = lj->ctx;
    const bool full = lj->symmetric && predIsSymmetric(pred);
    const unsigned grid = full ? unsigned(std::max<int64_t>(vl->ntiles, 1)) : gridFor(rows);
    lj->partials.ensure(ctx, int64_t(grid) * 2);
    double* partials = lj->partials.p;
    LJArgs a{at->pos.p, at->force.p, at->cap, 0, rows, 0, lj->dtable.p, int(lj->table.size()),
             int(lj->typeStride), lj->rcSqr, pred ? *pred : mrmd_pred2{}, pred && pred->op != MRMD_PAIR_ALWAYS};
    if (rows > 0)
    {
        if (full)
        {
            REQUIRE(vl->ntiles > 0, "the Verlet list has not been built");
            REQUIRE(!vl->banded, "the list holds only the rows within reach of the slab's weighted region (built "
                                 "inside the AdResS time loop); rebuild it for LennardJones");
            if (!fresh) vl->refreshSorted(at->pos.p);
            KTimer t(ctx, "lj_force");
            const TiledList tl{vl->nbr16.p, vl->cnt16.p, vl->strideS, vl->tiles.p, vl->bin.sorted.p,
                               vl->bin.sortedPos.p, vl->maxStaged, vl->sentinel,
                               vl->globalStage ? vl->gstage.p : nullptr, vl->gstageRecs};
            const Pre p1 = lj->table[0];
            const bool one = lj->table.size() == 1;
            const size_t lds = size_t(sentinelSlot(vl->maxStaged) + 1) *
                               (one ? Staged<true>::bytes : Staged<false>::bytes);
            const dim3 g(grid), b(vl->block);
            bool cap = false;
            for (const auto& p : lj->table) cap = cap || p.capDistSqr > 0.0;
            auto launch2 = [&](auto kernel, auto kernelGS) {
                if (vl->globalStage)
                    hipLaunchKernelGGL(kernelGS, g, b, 0, ctx->stream, a, tl, p1, partials);
                else
                    hipLaunchKernelGGL(kernel, g, b, lds, ctx->stream, a, tl, p1, partials);
            };
            const int variant = (a.usePred ? 4 : 0) | (one ? 2 : 0) | (cap ? 1 : 0);
            if (assign && variant == 2)
            {
                launch2(k_ljTiled<false, true, false, true, false>, k_ljTiled<false, true, false, true, true>);
                assigned = true;
            }
            else switch (variant)
            {
                case 0: launch2(k_ljTiled<false, false, false, false, false>, k_ljTiled<false, false, false, false, true>); break;
                case 1: launch2(k_ljTiled<false, false, true, false, false>, k_ljTiled<false, false, true, false, true>); break;
                case 2: launch2(k_ljTiled<false, true, false, false, false>, k_ljTiled<false, true, false, false, true>); break;
                case 3: launch2(k_ljTiled<false, true, true, false, false>, k_ljTiled<false, true, true, false, true>); break;
                case 4: launch2(k_ljTiled<true, false, false, false, false>, k_ljTiled<true, false, false, false, true>); break;
                case 5: launch2(k_ljTiled<true, false, true, false, false>, k_ljTiled<true, false, true, false, true>); break;
                case 6: launch2(k_ljTiled<true, true, false, false, false>, k_ljTiled<true, true, false, false, true>); break;
                default: launch2(k_ljTiled<true, true, true, false, false>, k_ljTiled<true, true, true, false, true>); break;
            }
        }
        else
        {
            vl->ensureHalf();
            a.stride = vl->stride;
            KTimer t(ctx, "lj_force_half");
            if (a.usePred)
                hipLaunchKernelGGL(k_ljHalf<true>, dim3(grid), dim3(BLOCK), 0, ctx->stream, a, vl->hnbr.p, vl->hcnt.p,
                                   partials);
            else
                hipLaunchKernelGGL(k_ljHalf<false>, dim3(grid), dim3(BLOCK), 0, ctx->stream, a, vl->hnbr.p, vl->hcnt.p,
                                   partials);
        }
        // summed when energy or virial is requested (most steps never ask)
        lj->nPartials = grid;
    }
    else
    {
        hipLaunchKernelGGL(k_zeroD, dim3(1), dim3(64), 0, ctx->stream, lj->ev.p, int64_t(2));
        lj->nPartials = 0;
    }
    checkLaunch();
    lj->evValid = false;
    return assigned;
}
}  // namespace

// ---------------------------------------------------------------------------------------
// Allocation-free apply (mrmd_lj_apply_noalloc): the per-workgroup energy / virial partials
// of the full-row kernel go to `scratch`, a file-scope device array of LJ_NOALLOC_MAX_TILES
// x 2 doubles (1 MiB) that each GPU holds until the process ends.  Host code never uses the
// array's name as a device address: its per-device address comes from
// hipGetSymbolAddress once per device, when a context is created on it.
namespace
{
constexpr int64_t LJ_NOALLOC_MAX_TILES = MRMD_LJ_NOALLOC_MAX_TILES;
static __device__ __align__(256) double scratch[2 * LJ_NOALLOC_MAX_TILES];
constexpr int NOALLOC_DEVICES = 64;
double* scratchDev[NOALLOC_DEVICES] = {};
std::mutex scratchMutex;

// sum of the partials into the caller's device memory and the table's own energy / virial
// (so that a later energy_virial() reads this call's values), one ordered pass
__global__ void k_sumPartialsTo(const double* partials, int64_t n, double* out, double* ev)
{
    __shared__ double red[BLOCK / WAVE];
    for (int c = 0; c < 2; ++c)
    {
        double s = 0.0;
        for (int64_t k = threadIdx.x; k < n; k += blockDim.x) s += partials[k * 2 + c];
        s = blockSum(s, red);
        if (threadIdx.x == 0)
        {
            out[c] = s;
            ev[c] = s;
        }
    }
}
}  // namespace

// called by mrmd_ctx_create with `device` current; a failed lookup leaves the device's
// entry empty, and mrmd_lj_apply_noalloc then fails with that message
bool mrmd_hip::ljNoAllocSetup(int device)
{
    std::lock_guard<std::mutex> lock(scratchMutex);
    if (device < 0 || device >= NOALLOC_DEVICES) return false;
    if (scratchDev[device]) return true;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
    {
        (void)hipGetLastError();
        return false;
    }
    scratchDev[device] = static_cast<double*>(p);
    return true;
}

extern "C" {

int mrmd_lj_apply_noalloc(mrmd_lj* lj, mrmd_atoms* at, mrmd_verlet_list* vl, double* ev_dev)
{
    MRMD_API_BEGIN(lj ? lj->ctx : nullptr)
    REQUIRE(lj && at && vl && ev_dev && at->ctx == lj->ctx && vl->ctx == lj->ctx, "invalid arguments");
    if (!lj->symmetric) fail(MRMD_ERR_UNSUP, "the allocation-free apply serves symmetric tables (the full-row kernel) only");
    REQUIRE(vl->begin == 0 && vl->end <= at->nlocal + at->nghost, "the list rows must start at record 0");
    REQUIRE(vl->nrows <= at->size, "the list was built for more records than the atoms hold");
    REQUIRE(!vl->banded, "the list holds only the rows within reach of the slab's weighted region (built inside "
                         "the AdResS time loop); rebuild it for LennardJones");
    REQUIRE(at->maxType * lj->typeStride + at->maxType < int64_t(lj->table.size()),
            "atom type %lld has no entry in the %lld-type table", (long long)at->maxType, (long long)lj->numTypes);
    if (vl->ntiles > LJ_NOALLOC_MAX_TILES)
        fail(MRMD_ERR_UNSUP, "the list has %lld tiles, the allocation-free apply's scratch holds %lld (use mrmd_lj_apply)",
             (long long)vl->ntiles, (long long)LJ_NOALLOC_MAX_TILES);
    mrmd_ctx* ctx = lj->ctx;
    double* part = ctx->device >= 0 && ctx->device < NOALLOC_DEVICES ? scratchDev[ctx->device] : nullptr;
    if (!part)
        fail(MRMD_ERR_HIP, "hipGetSymbolAddress gave no address for the allocation-free scratch on device %d",
             ctx->device);
    const int64_t rows = std::min<int64_t>(at->nlocal, vl->end);
    const bool fresh = vl->sortedFresh;
    vl->sortedFresh = false;
    if (rows > 0)
    {
        REQUIRE(vl->ntiles > 0, "the Verlet list has not been built");
        if (!fresh) vl->refreshSorted(at->pos.p);
        LJArgs a{at->pos.p, at->force.p, at->cap, 0, rows, 0, lj->dtable.p, int(lj->table.size()),
                 int(lj->typeStride), lj->rcSqr, mrmd_pred2{}, 0};
        const TiledList tl{vl->nbr16.p, vl->cnt16.p, vl->strideS, vl->tiles.p, vl->bin.sorted.p,
                           vl->bin.sortedPos.p, vl->maxStaged, vl->sentinel,
                               vl->globalStage ? vl->gstage.p : nullptr, vl->gstageRecs};
        const Pre p1 = lj->table[0];
        const bool one = lj->table.size() == 1;
        bool cap = false;
        for (const auto& q : lj->table) cap = cap || q.capDistSqr > 0.0;
        const size_t lds = size_t(sentinelSlot(vl->maxStaged) + 1) * (one ? Staged<true>::bytes : Staged<false>::bytes);
        const dim3 g(unsigned(vl->ntiles)), b(vl->block);
        auto launch2 = [&](auto kernel, auto kernelGS) {
            if (vl->globalStage)
                hipLaunchKernelGGL(kernelGS, g, b, 0, ctx->stream, a, tl, p1, part);
            else
                hipLaunchKernelGGL(kernel, g, b, lds, ctx->stream, a, tl, p1, part);
        };
        switch ((one ? 2 : 0) | (cap ? 1 : 0))
        {
            case 0: launch2(k_ljTiled<false, false, false, false, false>, k_ljTiled<false, false, false, false, true>); break;
            case 1: launch2(k_ljTiled<false, false, true, false, false>, k_ljTiled<false, false, true, false, true>); break;
            case 2: launch2(k_ljTiled<false, true, false, false, false>, k_ljTiled<false, true, false, false, true>); break;
            default: launch2(k_ljTiled<false, true, true, false, false>, k_ljTiled<false, true, true, false, true>); break;
        }
        hipLaunchKernelGGL(k_sumPartialsTo, dim3(1), dim3(BLOCK), 0, ctx->stream, part, vl->ntiles, ev_dev, lj->ev.p);
    }
    else
    {
        hipLaunchKernelGGL(k_zeroD, dim3(1), dim3(64), 0, ctx->stream, ev_dev, int64_t(2));
        hipLaunchKernelGGL(k_zeroD, dim3(1), dim3(64), 0, ctx->stream, lj->ev.p, int64_t(2));
    }
    checkLaunch();
    lj->nPartials = 0;
    lj->evValid = false;
    MRMD_API_END
}

int mrmd_lj_energy_virial(mrmd_lj* lj, double* e, double* w)
{
    MRMD_API_BEGIN(lj ? lj->ctx : nullptr)
    REQUIRE(lj, "null lj");
    if (!lj->evValid)
    {
        if (lj->nPartials > 0)
            hipLaunchKernelGGL(k_sumPartials, dim3(1), dim3(BLOCK), 0, lj->ctx->stream, lj->partials.p, lj->nPartials, 2,
                               lj->ev.p);
        checkLaunch();
        lj->nPartials = 0;
        HIPCHECK(hipMemcpyAsync(lj->ctx->pinned, lj->ev.p, 16, hipMemcpyDeviceToHost, lj->ctx->stream));
        sync(lj->ctx);
        lj->hE = lj->ctx->pinned[0];
        lj->hW = lj->ctx->pinned[1];
        lj->evValid = true;
    }
    if (e) *e = lj->hE;
    if (w) *w = lj->hW;
    MRMD_API_END
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// adjoint entry: ghost partners take the upstream gradient of their real atom
namespace mrmd_hip
{
const int32_t* ghostCorrDevice(const mrmd_ghost_layer* gl);
int64_t ghostCorrSize(const mrmd_ghost_layer* gl);
}

extern "C" int mrmd_lj_adjoint(mrmd_lj* lj, const mrmd_atoms* at, mrmd_verlet_list* vl, const mrmd_ghost_layer* gl,
                               const double* dLdF, double* dLdx, double* dLdeps, double* dLdsigma, int device)
{
    MRMD_API_BEGIN(lj ? lj->ctx : nullptr)
    REQUIRE(lj && at && vl && dLdF && dLdx && dLdeps && dLdsigma, "invalid arguments");
    REQUIRE(lj->symmetric, "the adjoint needs a symmetric parameter table");
    REQUIRE(vl->begin == 0, "the list rows must start at record 0");
    mrmd_ctx* ctx = lj->ctx;
    const int64_t nl = at->nlocal, nall = at->nlocal + at->nghost;
    REQUIRE(at->nghost == 0 || gl, "ghost records present: the ghost layer is required");
    REQUIRE(at->maxType * lj->typeStride + at->maxType < int64_t(lj->table.size()), "atom type outside the table");
    REQUIRE(vl->nrows <= nall, "the list was built for more records than the atoms hold");
    REQUIRE(at->nghost == 0 || mrmd_hip::ghostCorrSize(gl) >= nall, "the ghost layer does not cover these atoms");
    const int64_t rows = std::min<int64_t>(nl, vl->end);
    vl->ensureGlobal();
    const int nt2 = int(lj->table.size());
    const int64_t gstride = std::max<int64_t>(nall, 1);
    const int64_t xs = std::max<int64_t>(nl, 1);
    lj->adjG.ensure(ctx, 3 * gstride);
    lj->adjOut.ensure(ctx, 3 * xs);
    lj->adjStage.ensure(ctx, 3 * xs);
    const unsigned adjGrid = gridFor(std::max<int64_t>(rows, 1) * ADJ_LANES, ADJ_BLOCK);
    lj->adjPart.ensure(ctx, int64_t(adjGrid + 1) * 2 * nt2);
    double* const g = lj->adjG.p;
    double* const out = lj->adjOut.p;
    double* const stage = lj->adjStage.p;
    double* const pg = lj->adjPart.p;
    double* const pgSum = pg + int64_t(adjGrid) * 2 * nt2;
    // upstream: row-major n x 3 -> component-major with stride gstride (ghost records take
    // their real atom's)
    HIPCHECK(hipMemcpyAsync(stage, dLdF, size_t(3 * nl) * 8, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                            ctx->stream));
    if (nl) hipLaunchKernelGGL(k_rowMajorToSoA, dim3(gridFor(nl)), dim3(BLOCK), 0, ctx->stream, stage, g, nl, gstride);
    if (nall > nl)
        hipLaunchKernelGGL(k_expandGhostUpstream, dim3(gridFor(nall - nl)), dim3(BLOCK), 0, ctx->stream, g, gstride,
                           mrmd_hip::ghostCorrDevice(gl), nl, nall);
    hipLaunchKernelGGL(k_zeroD, dim3(1), dim3(BLOCK), 0, ctx->stream, pgSum, int64_t(2 * nt2));
    LJArgs a{at->pos.p, nullptr, at->cap, 0, rows, vl->stride, lj->dtable.p, nt2, int(lj->typeStride), lj->rcSqr,
             mrmd_pred2{}, 0};
    if (rows > 0)
    {
        KTimer t(ctx, "lj_adjoint");
        bool cap = false;
        for (const auto& p : lj->table) cap = cap || p.capDistSqr > 0.0;
        const dim3 grid(adjGrid), blk(ADJ_BLOCK);
        // sigma and epsilon of each pair from the parameters the table was built from
        const double* sd = lj->dParams.p;
        auto launch = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, grid, blk, 0, ctx->stream, a, vl->nbr.p, vl->cnt.p, g, gstride, sd, sd + nt2,
                               out, xs, pg);
        };
        if (nt2 <= 4)
            cap ? launch(k_ljAdjoint<true, true>) : launch(k_ljAdjoint<true, false>);
        else
            cap ? launch(k_ljAdjoint<false, true>) : launch(k_ljAdjoint<false, false>);
        hipLaunchKernelGGL(k_sumPartials, dim3(1), dim3(BLOCK), 0, ctx->stream, pg, int64_t(adjGrid), 2 * nt2, pgSum);
    }
    checkLaunch();
    // outputs: dL/dx back to row-major, the parameter sums to the host (one synchronisation)
    if (nl) hipLaunchKernelGGL(k_soaToRowMajor, dim3(gridFor(nl)), dim3(BLOCK), 0, ctx->stream, out, stage, nl, xs);
    HIPCHECK(hipMemcpyAsync(dLdx, stage, size_t(3 * nl) * 8, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                            ctx->stream));
    std::vector<double> hp(static_cast<size_t>(2 * nt2));
    HIPCHECK(hipMemcpyAsync(hp.data(), pgSum, size_t(2 * nt2) * 8, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    // symmetrise: the parameter of pair {a,b} is shared by table entries (a,b) and (b,a)
    const int64_t n = lj->numTypes, st = lj->typeStride;
    std::vector<double> ge(size_t(nt2), 0.0), gs(size_t(nt2), 0.0);
    for (int64_t x = 0; x < n; ++x)
        for (int64_t y = 0; y < n; ++y)
        {
            const int64_t i1 = x * st + y, i2 = y * st + x;
            if (i1 >= nt2 || i2 >= nt2) continue;
            ge[size_t(i1)] = i1 == i2 ? hp[size_t(i1)] : hp[size_t(i1)] + hp[size_t(i2)];
            gs[size_t(i1)] = i1 == i2 ? hp[size_t(nt2 + i1)] : hp[size_t(nt2 + i1)] + hp[size_t(nt2 + i2)];
        }
    if (device)
    {
        HIPCHECK(hipMemcpyAsync(dLdeps, ge.data(), size_t(nt2) * 8, hipMemcpyHostToDevice, ctx->stream));
        HIPCHECK(hipMemcpyAsync(dLdsigma, gs.data(), size_t(nt2) * 8, hipMemcpyHostToDevice, ctx->stream));
        sync(ctx);
    }
    else
    {
        std::memcpy(dLdeps, ge.data(), size_t(nt2) * 8);
        std::memcpy(dLdsigma, gs.data(), size_t(nt2) * 8);
    }
    MRMD_API_END
}
