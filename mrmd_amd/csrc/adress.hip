// AdResS LJ / ideal-gas coupling on the device: slab weighting, molecule update, the
// weighted molecule-pair force with drift and compensation terms, molecule -> atom force
// distribution, and the thermodynamic force (density sampling, smoothing, gradient,
// application).
//
// LJ_IdealGas (LJ_IdealGas.cpp:52-225) has two forms.  The weighting function is
// evaluated on ghost images, whose lambda need not equal that of their real molecule, so
// a pair must be evaluated with the lambda of the two copies the reference's half list
// pairs up.  The general form keeps the half-list algorithm (partner forces by FP64
// atomics) for molecules of several atoms and multi-type tables.  The tiled form, for
// lists whose every molecule is one atom and a single-type table, runs full rows from
// LDS-staged tiles with atomics only for partners that are not rows.  Only molecule
// pairs with at least one non-CG partner do pair work.  Density sampling accumulates into
// a workgroup-private LDS histogram before one global atomic per touched bin.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "nlist.hpp"
#include "physics.hpp"

using namespace mrmd_hip;

namespace
{
constexpr int COMP_BINS = 200;  // LJ_IdealGas::COMPENSATION_ENERGY_BINS (LJ_IdealGas.hpp:54)
constexpr int MAX_TF_CELLS = 4096;

__global__ void k_slabEval(mrmd_slab s, const double* x, int64_t n, double* out)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double l, m, g;
    slabEval(s, x[k], l, m, g);
    out[5 * k] = l;
    out[5 * k + 1] = m;
    out[5 * k + 2] = g;
    out[5 * k + 3] = 0.0;
    out[5 * k + 4] = 0.0;
}

__global__ void k_sphericalEval(mrmd_spherical s, const double* xyz, int64_t n, double* out)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double l, gx, gy, gz;
    sphericalEval(s, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], l, gx, gy, gz);
    out[4 * k] = l;
    out[4 * k + 1] = gx;
    out[4 * k + 2] = gy;
    out[4 * k + 3] = gz;
}

// UpdateMolecules::update (UpdateMolecules.hpp:24-70)
__global__ void k_updateMolecules(MolView m, AtomsView a, int64_t n, mrmd_slab s, int useSlab)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double X = 0.0, Y = 0.0, Z = 0.0;
    const int64_t start = m.offset[i], end = start + m.numAtoms[i];
    for (int64_t k = start; k < end; ++k)
    {
        const double4 p = a.pos[k];
        const double rm = a.relmass[k];
        X += p.x * rm;
        Y += p.y * rm;
        Z += p.z * rm;
    }
    double4 q = m.pos[i];
    q.x = X;
    q.y = Y;
    q.z = Z;
    m.pos[i] = q;
    if (useSlab)
    {
        double l, md, g;
        slabEval(s, X, l, md, g);
        m.lambda[i] = l;
        m.modLambda[i] = md;
        m.grad[i] = g;
        m.grad[m.cap + i] = 0.0;
        m.grad[2 * m.cap + i] = 0.0;
    }
}

// ContributeMoleculeForceToAtoms::update (ContributeMoleculeForceToAtoms.cpp:23-48)
__global__ void k_contribute(MolView m, AtomsView a, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double fx = m.force[i], fy = m.force[m.cap + i], fz = m.force[2 * m.cap + i];
    const int64_t start = m.offset[i], end = start + m.numAtoms[i];
    for (int64_t k = start; k < end; ++k)
    {
        const double rm = a.relmass[k];
        a.force[k] += rm * fx;
        a.force[a.cap + k] += rm * fy;
        a.force[2 * a.cap + k] += rm * fz;
    }
}

struct IGArgs
{
    MolView m;
    AtomsView a;
    const int32_t* nbr;
    const int32_t* cnt;
    int64_t stride;
    int64_t rows;
    const Pre* table;
    int nt;
    double rcSqr;
    int sampling;
    double compMin, compInvBin;
    double* comp;      // [COMP_BINS][nt]
    double* compCnt;   // [COMP_BINS][nt]
    const double* mean;
};

// LJ_IdealGas::operator() (LJ_IdealGas.cpp:52-225), half list over molecules
__global__ __launch_bounds__(BLOCK) void k_ljIdealGas(IGArgs g, double* partials)
{
    __shared__ Pre tab[MAX_TYPES2];
    __shared__ double red[BLOCK / WAVE];
    for (int t = threadIdx.x; t < g.nt * g.nt; t += blockDim.x) tab[t] = g.table[t];
    __syncthreads();
    const int64_t alpha = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    double energy = 0.0;
    if (alpha < g.rows)
    {
        const MolView& M = g.m;
        const AtomsView& A = g.a;
        double fA[3] = {0.0, 0.0, 0.0};
        const double mlA = M.modLambda[alpha];
        int64_t binA = -1;
        if (isHY(mlA)) binA = getBin(M.lambda[alpha], g.compMin, g.compInvBin, COMP_BINS);
        const double gA[3] = {M.grad[alpha], M.grad[M.cap + alpha], M.grad[2 * M.cap + alpha]};
        const int64_t sA = M.offset[alpha], eA = sA + M.numAtoms[alpha];
        const int nn = g.cnt[alpha];
        for (int n = 0; n < nn; ++n)
        {
            const int64_t beta = g.nbr[int64_t(n) * g.stride + alpha];
            const double mlB = M.modLambda[beta];
            if (isCG(mlA) && isCG(mlB)) continue;
            double fB[3] = {0.0, 0.0, 0.0};
            const double gB[3] = {M.grad[beta], M.grad[M.cap + beta], M.grad[2 * M.cap + beta]};
            const double w = 0.5 * (mlA + mlB);
            const bool hybridPair = isHY(mlA) || isHY(mlB);
            int64_t binB = -1;
            if (g.sampling && isHY(mlB)) binB = getBin(M.lambda[beta], g.compMin, g.compInvBin, COMP_BINS);
            const int64_t sB = M.offset[beta], eB = sB + M.numAtoms[beta];
            for (int64_t i = sA; i < eA; ++i)
            {
                const double4 xi = A.pos[i];
                const int ti = int(xi.w);
                double fi[3] = {0.0, 0.0, 0.0};
                for (int64_t j = sB; j < eB; ++j)
                {
                    const double4 xj = A.pos[j];
                    const double dx = xi.x - xj.x;
                    const double dy = xi.y - xj.y;
                    const double dz = xi.z - xj.z;
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 > g.rcSqr) continue;
                    const int tj = int(xj.w);
                    double ff, e;
                    forceEnergy(tab[ti * g.nt + tj], d2, ff, e);
                    const double ffw = ff * w;
                    fi[0] += dx * ffw;
                    fi[1] += dy * ffw;
                    fi[2] += dz * ffw;
                    atomicAddD(&A.force[j], -(dx * ffw));
                    atomicAddD(&A.force[A.cap + j], -(dy * ffw));
                    atomicAddD(&A.force[2 * A.cap + j], -(dz * ffw));
                    energy += e * w;
                    const double Vij = 0.5 * e;
                    if (hybridPair)
                    {
                        for (int d = 0; d < 3; ++d)
                        {
                            fA[d] += -Vij * gA[d];
                            fB[d] += -Vij * gB[d];
                        }
                        if (g.sampling)
                        {
                            if (isHY(mlA) && binA != -1) atomicAddD(&g.comp[binA * g.nt + ti], Vij);
                            if (isHY(mlB) && binB != -1) atomicAddD(&g.comp[binB * g.nt + tj], Vij);
                        }
                    }
                }
                atomicAddD(&A.force[i], fi[0]);
                atomicAddD(&A.force[A.cap + i], fi[1]);
                atomicAddD(&A.force[2 * A.cap + i], fi[2]);
            }
            if (fB[0] != 0.0 || fB[1] != 0.0 || fB[2] != 0.0)
                for (int d = 0; d < 3; ++d) atomicAddD(&M.force[d * M.cap + beta], fB[d]);
        }
        if (g.sampling && isHY(mlA) && binA != -1)
            for (int64_t k = sA; k < eA; ++k) atomicAddD(&g.compCnt[binA * g.nt + int(A.pos[k].w)], 1.0);
        if (isHY(mlA) && binA != -1)
        {
            const double mc = g.mean[binA * g.nt + int(A.pos[sA].w)];
            for (int d = 0; d < 3; ++d) fA[d] += mc * gA[d];
        }
        for (int d = 0; d < 3; ++d) atomicAddD(&M.force[d * M.cap + alpha], fA[d]);
    }
    energy = blockSum(energy, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = energy;
}

// ---- tiled form for molecule lists whose every record is one atom ----
//
// The half list holds each physical pair once: either as (row, partner) in the row's list
// or, when the partner is a non-row record (a ghost image), possibly only as the mirror
// pair (partner's real, row's image) in the real's list -- the position-lexicographic
// discriminator picks exactly one of the two images.  A row of the full tiled list
// therefore takes
//   * every row partner (that partner's own row adds the other side), and
//   * a non-row partner only when the entry carries the half-list flag; the partner's
//     side (atom force, molecule drift force, compensation sample) then goes to the
//     non-row record through FP64 atomics, exactly as the reference deposits it before
//     the ghost fold.
// Unflagged non-row partners are skipped: the mirror pair deposits both sides.  Every
// pair is thus evaluated with the lambda of the very copies the reference pairs up, for
// any slab placement, and the row's own sums need no atomics.
struct IGTiled
{
    const uint16_t* nbr;
    const int32_t* cnt;
    int64_t strideS;
    const TileInfo* tiles;
    const int32_t* sorted;
    const double4* igPos;  ///< sorted slot -> {atom x, y, z, lambda_mod}, sign bit set on non-rows
    int64_t rowsEnd;       ///< molecules [0, rowsEnd) are rows
    int queueCap;          ///< queued non-row partner pairs per workgroup (<= IG_QUEUE)
    int32_t maxStaged;     ///< the sentinel record sits at sentinelSlot(maxStaged)
    double sentinel;       ///< at (sentinel, sentinel, sentinel)
    const int8_t* skip;    ///< per tile: 1 = no record of its neighbourhood can be weighted (or null)
};

// sorted-slot record of each molecule's single atom and its modulated lambda (in [0, 1];
// stored negated, -0.0 included, for molecules that are not rows)
__global__ void k_igGather(MolView m, const double4* __restrict__ atomPos, const int32_t* __restrict__ sorted,
                           int64_t n, int64_t rowsEnd, double4* __restrict__ out)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int32_t k = sorted[s];
    const double4 p = atomPos[m.offset[k]];
    const double l = fabs(m.modLambda[k]);
    out[s] = make_double4(p.x, p.y, p.z, k < rowsEnd ? l : -l);
}

__device__ inline int stagedSlot(const TileInfo& ti, int q)
{
    int r = 0;
    while (r + 1 < ti.nr && ti.base[r + 1] <= q) ++r;
    return ti.begin[r] + (q - ti.base[r]);
}

// pair terms of one chunk of eight entries (single-type table).  Returns the mask of
// entries that pair the row with a flagged non-row partner.
template <bool CAP>
__device__ inline uint32_t igChunk(const IGArgs& g, const Pre& p, const double4& pi, const double4* st,
                                   const uint4 cur, double& fx, double& fy, double& fz, double& ew, double& v)
{
#pragma clang fp contract(fast)
    const uint32_t wd[4] = {cur.x, cur.y, cur.z, cur.w};
    double dx[CHUNK], dy[CHUNK], dz[CHUNK], d2[CHUNK], lb[CHUNK];
    uint32_t flag = 0;
#pragma unroll
    for (int u = 0; u < CHUNK; ++u)
    {
        const uint32_t e = (wd[u >> 1] >> ((u & 1) * 16)) & 0xffffu;
        const double4 pj = st[e & 0x7fffu];
        dx[u] = pi.x - pj.x;
        dy[u] = pi.y - pj.y;
        dz[u] = pi.z - pj.z;
        lb[u] = pj.w;
        flag |= (e >> 15) << u;
    }
#pragma unroll
    for (int u = 0; u < CHUNK; ++u) d2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
    double frac2[CHUNK];
#pragma unroll
    for (int u = 0; u < CHUNK; ++u) frac2[u] = recipNewton(d2[u]);
    const bool hyA = isHY(pi.w);
    uint32_t remote = 0;
    auto accumulate = [&](int u, double ff, double en) {
        const bool nonRow = signbit(lb[u]);
        const double lam = fabs(lb[u]);
        // beyond rc (padding entries point at the sentinel record), both CG
        // (LJ_IdealGas.cpp:95), unflagged non-row
        const bool in = d2[u] <= g.rcSqr && d2[u] > 0.0 && !(isCG(pi.w) && isCG(lam)) &&
                        (!nonRow || ((flag >> u) & 1u));
        const double w = 0.5 * (pi.w + lam);
        const double ffw = in ? ff * w : 0.0;
        fx += dx[u] * ffw;
        fy += dy[u] * ffw;
        fz += dz[u] * ffw;
        const double ewu = en * w;
        ew += in ? (nonRow ? ewu : 0.5 * ewu) : 0.0;
        v += (in && (hyA || isHY(lam))) ? 0.5 * en : 0.0;
        remote |= uint32_t(in && nonRow) << u;
    };
    if constexpr (!CAP)
    {
#pragma unroll
        for (int u = 0; u < CHUNK; ++u)
        {
            const double frac6 = frac2[u] * frac2[u] * frac2[u];
            accumulate(u, frac6 * (p.ff1 * frac6 - p.ff2) * frac2[u], frac6 * (p.ef1 * frac6 - p.ef2) - p.shift);
        }
        return remote;
    }
    double ffv[CHUNK], env[CHUNK];
    bool capped = false;
#pragma unroll
    for (int u = 0; u < CHUNK; ++u)
    {
        const double frac6 = frac2[u] * frac2[u] * frac2[u];
        ffv[u] = frac6 * (p.ff1 * frac6 - p.ff2) * frac2[u];
        env[u] = frac6 * (p.ef1 * frac6 - p.ef2) - p.shift;
        capped = capped || d2[u] < p.capDistSqr;
    }
    // pairs inside the capping distance are rare: the capped form (LennardJones.hpp:68-75)
    // runs, by selects over the chunk, only where some lane of the wave has one
    if (__any(capped))
    {
        // keeps the block a branch: the compiler would otherwise evaluate it for every chunk
        asm volatile("" ::: "memory");
#pragma unroll
        for (int u = 0; u < CHUNK; ++u)
        {
            const double dist = sqrt(d2[u]);
            const bool c = d2[u] < p.capDistSqr;
            ffv[u] = c ? p.capCoeff / dist : ffv[u];
            env[u] = c ? p.eCap - (dist - p.capDist) * p.capCoeff - p.shift : env[u];
        }
    }
#pragma unroll
    for (int u = 0; u < CHUNK; ++u) accumulate(u, ffv[u], env[u]);
    return remote;
}

// the partner side of one (row, flagged non-row) pair: atom force, drift on the partner
// molecule and its compensation sample (LJ_IdealGas.cpp:120-190), by atomics
template <bool CAP>
__device__ inline void igRemotePair(const IGArgs& g, const IGTiled& tl, const TileInfo& ti, const Pre& p,
                                    const double4& pi, const double4* st, int q)
{
    const MolView& M = g.m;
    const AtomsView& A = g.a;
    const double4 pj = st[q];
    const double lam = fabs(pj.w);
    const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    double ff, en;
    if (CAP && d2 < p.capDistSqr)
    {
        const double dist = sqrt(d2);
        ff = p.capCoeff / dist;
        en = p.eCap - (dist - p.capDist) * p.capCoeff - p.shift;
    }
    else
    {
        const double frac2 = 1.0 / d2;
        const double frac6 = frac2 * frac2 * frac2;
        ff = frac6 * (p.ff1 * frac6 - p.ff2) * frac2;
        en = frac6 * (p.ef1 * frac6 - p.ef2) - p.shift;
    }
    const double ffw = ff * (0.5 * (pi.w + lam));
    const int32_t k = tl.sorted[stagedSlot(ti, q)];
    const int64_t b = M.offset[k];
    atomicAddD(&A.force[b], -(dx * ffw));
    atomicAddD(&A.force[A.cap + b], -(dy * ffw));
    atomicAddD(&A.force[2 * A.cap + b], -(dz * ffw));
    if (isHY(pi.w) || isHY(lam))
    {
        const double Vij = 0.5 * en;
        for (int d = 0; d < 3; ++d)
        {
            const double gB = M.grad[d * M.cap + k];
            if (gB != 0.0) atomicAddD(&M.force[d * M.cap + k], -Vij * gB);
        }
        if (g.sampling && isHY(lam))
        {
            const int64_t binB = getBin(M.lambda[k], g.compMin, g.compInvBin, COMP_BINS);
            if (binB != -1) atomicAddD(&g.comp[binB], Vij);
        }
    }
}

// the (row, flagged non-row) pairs of one row's chunks from ch0 on, by a second pass over
// its entries (used once the workgroup's queue of such pairs is full)
template <bool CAP>
__device__ void igRemoteRow(const IGArgs& g, const IGTiled& tl, const TileInfo& ti, const Pre& p,
                            const double4& pi, const double4* st, int s, int c, int ch0)
{
    const uint4* chunk = reinterpret_cast<const uint4*>(tl.nbr) + s;
    for (int ch = ch0; ch * CHUNK < c; ++ch)
    {
        const uint4 cur = chunk[int64_t(ch) * tl.strideS];
        const uint32_t wd[4] = {cur.x, cur.y, cur.z, cur.w};
        for (int u = 0; u < CHUNK && ch * CHUNK + u < c; ++u)
        {
            const uint32_t e = (wd[u >> 1] >> ((u & 1) * 16)) & 0xffffu;
            const int q = int(e & 0x7fffu);
            const double4 pj = st[q];
            if (!(e & 0x8000u) || !signbit(pj.w)) continue;
            const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 <= g.rcSqr && d2 > 0.0 && !(isCG(pi.w) && isCG(fabs(pj.w))))
                igRemotePair<CAP>(g, tl, ti, p, pi, st, q);
        }
    }
}

constexpr int IG_QUEUE = 4096;  // (row, partner) pairs with a non-row partner, per workgroup
constexpr uint32_t IG_SKIP = 0xffffffffu;  // a queue slot reserved by a reservation that failed

template <bool CAP>
__global__ __launch_bounds__(BLOCK) void k_igTiled(IGArgs g, IGTiled tl, Pre p, double* partials)
{
    extern __shared__ double4 st[];
    __shared__ TileInfo ti;
    __shared__ double red[BLOCK / WAVE];
    __shared__ uint32_t queue[IG_QUEUE];  // LDS slot of the row << 16 | LDS slot of the partner
    __shared__ int queued;
    // consecutive tiles go to different XCDs: the pair work sits in the tiles of the
    // atomistic and hybrid slab, one contiguous range of tile indices, which an
    // XCD-contiguous remap would hand to two or three XCDs only
    const unsigned blk = blockIdx.x;
    // a tile that the banded list build found out of reach of the weighted region
    if (tl.skip && tl.skip[blk])
    {
        if (threadIdx.x == 0) partials[blk] = 0.0;
        return;
    }
    if (threadIdx.x == 0)
    {
        ti = tl.tiles[blk];
        queued = 0;
    }
    __syncthreads();
    // the first row's loads overlap the staging
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
        pi = tl.igPos[s];
        cur = reinterpret_cast<const uint4*>(tl.nbr)[s];
    }
    stageTile<false>(ti, tl.igPos, st);
    putSentinel<false>(st, tl.maxStaged, tl.sentinel, 0.0);
    __syncthreads();
    // a neighbourhood of coarse-grained molecules only (lambda_mod = 0 everywhere) has no
    // pair work (LJ_IdealGas.cpp:95), no drift and no compensation term
    int busy = 0;
    for (int q = threadIdx.x; q < ti.staged; q += blockDim.x) busy |= st[q].w != 0.0;
    if (!__syncthreads_or(busy))
    {
        if (threadIdx.x == 0) partials[blk] = 0.0;
        return;
    }
    const MolView& M = g.m;
    const AtomsView& A = g.a;
    double eTile = 0.0;
    for (int t = threadIdx.x; t < ti.nrows; t += blockDim.x)
    {
        int self = 0;
        if (t != int(threadIdx.x))
        {
            s = tileRow(ti, t, self);
            c = tl.cnt[s];
            i = tl.sorted[s];
            pi = tl.igPos[s];
            cur = reinterpret_cast<const uint4*>(tl.nbr)[s];
        }
        else
            tileRow(ti, t, self);
        if (c == 0 || i >= tl.rowsEnd) continue;
        // the row's own record sits in its column's staged range
        const uint32_t ownSlot = uint32_t(ldsSlot(ti, ti.selfRange[self], s));
        double fx = 0.0, fy = 0.0, fz = 0.0, v = 0.0;
        const uint4* chunk = reinterpret_cast<const uint4*>(tl.nbr) + s;
        const int nch = (c + CHUNK - 1) / CHUNK;
        int overflow = -1;  // first chunk whose partner sides did not fit into the queue
        for (int ch = 0; ch < nch; ++ch)
        {
            uint4 nxt = cur;
            if (ch + 1 < nch) nxt = chunk[int64_t(ch + 1) * tl.strideS];
            uint32_t remote = igChunk<CAP>(g, p, pi, st, cur, fx, fy, fz, eTile, v);
            if (remote && overflow < 0)
            {
                // the partner sides go to the workgroup's queue, shared out after the rows
                const int n = __builtin_popcount(remote);
                int at = atomicAdd(&queued, n);
                if (at + n > tl.queueCap)
                {
                    // every later reservation fails too; slots of this one that exist are
                    // marked so that the queue pass skips them
                    overflow = ch;
                    for (int k = at; k < tl.queueCap; ++k) queue[k] = IG_SKIP;
                }
                else
                    while (remote)
                    {
                        const int b = __builtin_ctz(remote);
                        remote &= remote - 1;
                        const uint32_t w2 = b < 4 ? (b < 2 ? cur.x : cur.y) : (b < 6 ? cur.z : cur.w);
                        queue[at++] = (ownSlot << 16) | ((w2 >> ((b & 1) * 16)) & 0x7fffu);
                    }
            }
            cur = nxt;
        }
        // partner sides that did not fit are deposited by the row itself
        if (overflow >= 0) igRemoteRow<CAP>(g, tl, ti, p, pi, st, s, c, overflow);
        const int64_t a = M.offset[i];
        A.force[a] += fx;
        A.force[A.cap + a] += fy;
        A.force[2 * A.cap + a] += fz;
        // drift -1/2 sum_j V_ij grad(lambda_i) over hybrid pairs and the mean compensation
        // energy (LJ_IdealGas.cpp:150-170, 208-224)
        const double gA[3] = {M.grad[i], M.grad[M.cap + i], M.grad[2 * M.cap + i]};
        int64_t binA = -1;
        if (isHY(pi.w)) binA = getBin(M.lambda[i], g.compMin, g.compInvBin, COMP_BINS);
        const double mc = binA != -1 ? g.mean[binA] : 0.0;
        for (int d = 0; d < 3; ++d) M.force[d * M.cap + i] += -v * gA[d] + mc * gA[d];
        if (g.sampling && binA != -1)
        {
            atomicAddD(&g.comp[binA], v);
            atomicAddD(&g.compCnt[binA], 1.0);
        }
    }
    __syncthreads();
    const int nq = min(queued, tl.queueCap);
    for (int k = threadIdx.x; k < nq; k += blockDim.x)
    {
        const uint32_t item = queue[k];
        if (item == IG_SKIP) continue;
        igRemotePair<CAP>(g, tl, ti, p, st[item >> 16], st, int(item & 0x7fffu));
    }
    eTile = blockSum(eTile, red);
    if (threadIdx.x == 0) partials[blk] = eTile;
}

// updateMeanCompensationEnergy (LJ_IdealGas.cpp:21-50)
__global__ void k_updateMeanComp(double* comp, double* cnt, double* mean, int64_t n, double factor)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (cnt[k] < 0.5) return;
    const double e = comp[k] / cnt[k];
    mean[k] = (factor * mean[k] + e) / (factor + 1.0);
    comp[k] = 0.0;
    cnt[k] = 0.0;
}

// host (if not null): pinned host memory that receives the sum too, written by the kernel
__global__ void k_sumOrdered(const double* partials, int64_t n, double* out, double* host)
{
    __shared__ double red[BLOCK / WAVE];
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) s += partials[k];
    s = blockSum(s, red);
    if (threadIdx.x == 0)
    {
        *out = s;
        if (host)
        {
            *host = s;
            __threadfence_system();
        }
    }
}

__global__ void k_zeroD(double* p, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0;
}

// ---------------------------------------------------------------------------------------
// thermodynamic force
// getAxialDensityProfile (analysis/AxialDensityProfile.cpp:21-51) accumulated into the
// running profile: workgroup-private LDS histogram, then one global atomic per bin
__global__ __launch_bounds__(BLOCK) void k_tfSample(const double4* pos, int64_t n, HistD h, double* density)
{
    __shared__ double loc[MAX_TF_CELLS];
    const int64_t cells = h.numBins * h.numHist;
    for (int64_t c = threadIdx.x; c < cells; c += blockDim.x) loc[c] = 0.0;
    __syncthreads();
    // four records in flight per lane, then their LDS increments
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i0 = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * stride)
    {
        double4 p[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const int64_t i = i0 + u * stride;
            p[u] = i < n ? pos[i] : make_double4(h.min - 1.0, 0.0, 0.0, 0.0);  // below min: no bin
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
        {
            const int64_t bin = getBin(p[u].x, h.min, h.invBinSize, h.numBins);
            if (bin != -1) atomicAdd(&loc[bin * h.numHist + int64_t(p[u].w)], 1.0);
        }
    }
    __syncthreads();
    for (int64_t c = threadIdx.x; c < cells; c += blockDim.x)
        if (loc[c] != 0.0) atomicAddD(&density[c], loc[c]);
}

// ThermodynamicForce::update_if (ThermodynamicForce.hpp:124-152), one workgroup
__global__ __launch_bounds__(BLOCK) void k_tfUpdate(double* density, double* force, HistD hd, double norm,
                                                    const double* forceFactor, double sigma, int64_t delta,
                                                    int sym, int periodic, mrmd_pred1 pred, int usePred)
{
    __shared__ double dens[MAX_TF_CELLS];
    __shared__ double smooth[MAX_TF_CELLS];
    const int64_t nb = hd.numBins, nh = hd.numHist, cells = nb * nh;
    for (int64_t c = threadIdx.x; c < cells; c += blockDim.x) dens[c] = density[c];
    __syncthreads();
    if (sym)
    {
        const int64_t maxIdx = nb - 1;
        for (int64_t c = threadIdx.x; c < (nb / 2) * nh; c += blockDim.x)
        {
            const int64_t i = c / nh, h = c % nh;
            const double v = 0.5 * (dens[i * nh + h] + dens[(maxIdx - i) * nh + h]);
            dens[i * nh + h] = v;
            dens[(maxIdx - i) * nh + h] = v;
        }
        __syncthreads();
    }
    for (int64_t c = threadIdx.x; c < cells; c += blockDim.x) dens[c] *= norm;
    __syncthreads();
    for (int64_t c = threadIdx.x; c < cells; c += blockDim.x)
        smooth[c] = smoothAt(dens, hd, c / nh, c % nh, sigma, delta, periodic);
    __syncthreads();
    for (int64_t c = threadIdx.x; c < cells; c += blockDim.x)
    {
        const int64_t b = c / nh, h = c % nh;
        double gr = gradientAt(smooth, hd, b, h, periodic) * forceFactor[h];
        const double x = hd.min + (double(b) + 0.5) * hd.binSize;
        if (usePred && !evalPred1(pred, x, x, x)) gr = 0.0;
        force[c] -= gr;
        density[c] = 0.0;
    }
}

// ThermodynamicForce::apply_if (ThermodynamicForce.hpp:98-122)
template <bool PRED>
__global__ void k_tfApply(AtomsView a, int64_t n, const double* force, HistD hd, mrmd_pred1 pred)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = a.pos[i];
    if (PRED && !evalPred1(pred, p.x, p.y, p.z)) return;
    const int64_t bin = getBin(p.x, hd.min, hd.invBinSize, hd.numBins);
    if (bin != -1) a.force[i] += force[bin * hd.numHist + int64_t(p.w)];
}

// ThermodynamicForce::applyInterpolated_if (ThermodynamicForce.hpp:154-216)
template <bool PRED>
__global__ void k_tfApplyInterp(AtomsView a, int64_t n, const double* force, HistD hd, mrmd_pred1 pred)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = a.pos[i];
    if (PRED && !evalPred1(pred, p.x, p.y, p.z)) return;
    const double x = p.x;
    const int64_t bin = getBin(x, hd.min, hd.invBinSize, hd.numBins);
    if (bin == -1) return;
    const int64_t t = int64_t(p.w);
    const double binStart = hd.min + double(bin) * hd.binSize;
    const double frac = (x - binStart) * hd.invBinSize;
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
    if (l >= 0 && r < hd.numBins)
    {
        const double L = force[l * hd.numHist + t], R = force[r * hd.numHist + t];
        a.force[i] += L + (R - L) * fac;
    }
    else
        a.force[i] += force[bin * hd.numHist + t];
}

__global__ void k_histGradient(const double* in, double* out, HistD hd, int periodic)
{
    for (int64_t c = threadIdx.x; c < hd.numBins * hd.numHist; c += blockDim.x)
        out[c] = gradientAt(in, hd, c / hd.numHist, c % hd.numHist, periodic);
}
__global__ void k_histSmoothen(const double* in, double* out, HistD hd, double sigma, int64_t delta, int periodic)
{
    for (int64_t c = threadIdx.x; c < hd.numBins * hd.numHist; c += blockDim.x)
        out[c] = smoothAt(in, hd, c / hd.numHist, c % hd.numHist, sigma, delta, periodic);
}

HistD makeHist(double mn, double mx, int64_t nb, int64_t nh)
{
    HistD h;
    h.min = mn;
    h.numBins = nb;
    h.numHist = nh;
    h.binSize = (mx - mn) / double(nb);
    h.invBinSize = 1.0 / h.binSize;
    return h;
}
}  // namespace

struct mrmd_lj_idealgas
{
    mrmd_hip::CtxPtr ctx;
    int64_t numTypes = 1;
    std::vector<Pre> table;
    double rcSqr = 0.0;
    DBuf<Pre> dtable;
    DBuf<double> hist;  ///< comp | count | mean, each COMP_BINS x numTypes
    DBuf<double4> igPos;  ///< tiled form: sorted-slot atom records
    /// set when a time loop's fused pass has already written igPos for the next run (for the
    /// list `igPosList` at its current build); the run then skips its gather
    bool igPosFresh = false;
    const mrmd_verlet_list* igPosList = nullptr;
    int queueCap = IG_QUEUE;
    int64_t runCounter = 0;
    int64_t samplingInterval = 200;
    int64_t updateInterval = 20000;
};

struct mrmd_thermo_force
{
    mrmd_hip::CtxPtr ctx;
    HistD hd{};
    double max = 0.0;
    double binVolume = 1.0;
    int64_t samples = 0;
    bool sym = false, periodic = false;
    std::vector<double> forceFactor;
    DBuf<double> force, density, factor;
};

extern "C" {

int mrmd_slab_eval(mrmd_ctx* c, const mrmd_slab* s, const double* x, int64_t n, double* out5)
{
    MRMD_API_BEGIN(c)
    REQUIRE(s && x && out5 && n >= 0, "invalid arguments");
    if (n == 0) return MRMD_OK;
    DBuf<double> dx, dout;
    dx.reserve(c, n);
    dout.reserve(c, 5 * n);
    HIPCHECK(hipMemcpyAsync(dx.p, x, size_t(n) * 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_slabEval, dim3(gridFor(n)), dim3(BLOCK), 0, c->stream, *s, dx.p, n, dout.p);
    checkLaunch();
    HIPCHECK(hipMemcpyAsync(out5, dout.p, size_t(5 * n) * 8, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    MRMD_API_END
}

int mrmd_spherical_eval(mrmd_ctx* c, const mrmd_spherical* s, const double* xyz, int64_t n, double* out4)
{
    MRMD_API_BEGIN(c)
    REQUIRE(s && xyz && out4 && n >= 0, "invalid arguments");
    REQUIRE(s->hybrid_diameter > 0.0, "the hybrid shell needs a positive width");
    if (n == 0) return MRMD_OK;
    DBuf<double> dx, dout;
    dx.reserve(c, 3 * n);
    dout.reserve(c, 4 * n);
    HIPCHECK(hipMemcpyAsync(dx.p, xyz, size_t(3 * n) * 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_sphericalEval, dim3(gridFor(n)), dim3(BLOCK), 0, c->stream, *s, dx.p, n, dout.p);
    checkLaunch();
    HIPCHECK(hipMemcpyAsync(out4, dout.p, size_t(4 * n) * 8, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    MRMD_API_END
}

int mrmd_molecules_update(mrmd_molecules* m, const mrmd_atoms* a, const mrmd_slab* s)
{
    MRMD_API_BEGIN(m ? m->ctx : nullptr)
    REQUIRE(m && a, "invalid arguments");
    const int64_t n = m->nlocal + m->nghost;
    if (n == 0) return MRMD_OK;
    KTimer t(m->ctx, "update_molecules");
    hipLaunchKernelGGL(k_updateMolecules, dim3(gridFor(n)), dim3(BLOCK), 0, m->ctx->stream, m->view(), a->view(), n,
                       s ? *s : mrmd_slab{}, s ? 1 : 0);
    checkLaunch();
    MRMD_API_END
}

int mrmd_molecules_contribute_force(const mrmd_molecules* mc, mrmd_atoms* a)
{
    auto* m = const_cast<mrmd_molecules*>(mc);
    MRMD_API_BEGIN(m ? m->ctx : nullptr)
    REQUIRE(m && a, "invalid arguments");
    const int64_t n = m->nlocal + m->nghost;
    if (n == 0) return MRMD_OK;
    KTimer t(m->ctx, "contribute_molecule_force");
    hipLaunchKernelGGL(k_contribute, dim3(gridFor(n)), dim3(BLOCK), 0, m->ctx->stream, m->view(), a->view(), n);
    checkLaunch();
    MRMD_API_END
}

int mrmd_ljig_create(mrmd_ctx* c, const double* cap, const double* rc, const double* sigma, const double* eps,
                     int64_t numTypes, int shifted, mrmd_lj_idealgas** out)
{
    MRMD_API_BEGIN(c)
    REQUIRE(out && cap && rc && sigma && eps && numTypes >= 1, "invalid arguments");
    REQUIRE(numTypes * numTypes <= MAX_TYPES2, "at most %d type pairs", MAX_TYPES2);
    auto* s = new mrmd_lj_idealgas();
    s->ctx = c;
    s->numTypes = numTypes;
    const int64_t nt2 = numTypes * numTypes;
    buildTable(s->table, cap, rc, sigma, eps, nt2, shifted != 0);
    double rcMax = rc[0];
    for (int64_t k = 0; k < nt2; ++k) rcMax = std::max(rcMax, rc[k]);
    s->rcSqr = rcMax * rcMax;
    try
    {
        s->dtable.reserve(c, nt2);
        HIPCHECK(hipMemcpyAsync(s->dtable.p, s->table.data(), size_t(nt2) * sizeof(Pre), hipMemcpyHostToDevice, c->stream));
        s->hist.reserve(c, 3 * COMP_BINS * numTypes);
        hipLaunchKernelGGL(k_zeroD, dim3(gridFor(3 * COMP_BINS * numTypes)), dim3(BLOCK), 0, c->stream, s->hist.p,
                           int64_t(3 * COMP_BINS * numTypes));
        checkLaunch();
        sync(c);
    }
    catch (...)
    {
        delete s;
        throw;
    }
    *out = s;
    MRMD_API_END
}

int mrmd_ljig_destroy(mrmd_lj_idealgas* s)
{
    if (!s) return MRMD_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
    return MRMD_OK;
}

int mrmd_ljig_set_intervals(mrmd_lj_idealgas* s, int64_t sampling, int64_t update)
{
    MRMD_API_BEGIN(s ? s->ctx : nullptr)
    REQUIRE(s && sampling > 0 && update > 0, "intervals must be positive");
    s->samplingInterval = sampling;
    s->updateInterval = update;
    MRMD_API_END
}

int mrmd_ljig_set_queue(mrmd_lj_idealgas* s, int capacity)
{
    MRMD_API_BEGIN(s ? s->ctx : nullptr)
    REQUIRE(s && capacity >= 0 && capacity <= IG_QUEUE, "queue capacity must be in [0, %d]", IG_QUEUE);
    s->queueCap = capacity;
    MRMD_API_END
}

int mrmd_ljig_run(mrmd_lj_idealgas* s, mrmd_molecules* m, mrmd_verlet_list* vl, mrmd_atoms* a, double* energy)
{
    MRMD_API_BEGIN(s ? s->ctx : nullptr)
    REQUIRE(s && m && vl && a, "invalid arguments");
    REQUIRE(vl->begin == 0, "the molecule list rows must start at record 0");
    REQUIRE(a->maxType < s->numTypes, "atom type %lld outside the %lld-type table", (long long)a->maxType,
            (long long)s->numTypes);
    REQUIRE(vl->nrows <= m->nlocal + m->nghost, "the list was built for more molecules than exist");
    mrmd_ctx* ctx = s->ctx;
    const int64_t rows = std::min<int64_t>(m->nlocal, vl->end);
    // the tiled form serves single-type tables over lists whose molecules are single atoms
    // (checked when the list was built, valid while the molecules' layout is unchanged)
    const bool tiled = s->numTypes == 1 && vl->monoOf == m && vl->monoVersion == m->layoutVersion &&
                       vl->ntiles > 0 && vl->nrows == m->nlocal + m->nghost && !vl->globalStage;
    if (!tiled) vl->ensureHalf();
    const unsigned grid = tiled ? unsigned(vl->ntiles) : gridFor(rows);
    double* part = static_cast<double*>(ctxScratch(ctx, 1, size_t(grid + 1) * sizeof(double)));
    // the energy's host copy: written by the summing kernel, 0.0 when nothing runs
    double* host = nullptr;
    if (energy)
    {
        ctx->pinned[0] = 0.0;
        host = ctx->pinned;
    }
    const int64_t nh = COMP_BINS * s->numTypes;
    const double compBin = (1.0 - 0.0) / double(COMP_BINS);
    IGArgs g{m->view(), a->view(), vl->hnbr.p, vl->hcnt.p, vl->stride, rows, s->dtable.p, int(s->numTypes), s->rcSqr,
             (s->runCounter % s->samplingInterval) == 0 ? 1 : 0, 0.0, 1.0 / compBin, s->hist.p, s->hist.p + nh,
             s->hist.p + 2 * nh};
    if (rows > 0 && tiled)
    {
        KTimer t(ctx, "lj_idealgas_tiled");
        if (!(s->igPosFresh && s->igPosList == vl))
        {
            s->igPos.ensure(ctx, vl->nrows);
            hipLaunchKernelGGL(k_igGather, dim3(gridFor(vl->nrows)), dim3(BLOCK), 0, ctx->stream, m->view(), a->pos.p,
                               vl->bin.sorted.p, vl->nrows, rows, s->igPos.p);
        }
        const IGTiled tl{vl->nbr16.p, vl->cnt16.p, vl->strideS, vl->tiles.p, vl->bin.sorted.p, s->igPos.p, rows,
                         s->queueCap, vl->maxStaged, vl->sentinel, vl->banded ? vl->tileSkip.p : nullptr};
        const Pre p = s->table[0];
        const size_t lds = size_t(sentinelSlot(vl->maxStaged) + 1) * sizeof(double4);
        if (p.capDistSqr > 0.0)
            hipLaunchKernelGGL(k_igTiled<true>, dim3(grid), dim3(vl->block), lds, ctx->stream, g, tl, p, part);
        else
            hipLaunchKernelGGL(k_igTiled<false>, dim3(grid), dim3(vl->block), lds, ctx->stream, g, tl, p, part);
        hipLaunchKernelGGL(k_sumOrdered, dim3(1), dim3(BLOCK), 0, ctx->stream, part, int64_t(grid), part + grid,
                           host);
    }
    else if (rows > 0)
    {
        KTimer t(ctx, "lj_idealgas");
        hipLaunchKernelGGL(k_ljIdealGas, dim3(grid), dim3(BLOCK), 0, ctx->stream, g, part);
        hipLaunchKernelGGL(k_sumOrdered, dim3(1), dim3(BLOCK), 0, ctx->stream, part, int64_t(grid), part + grid,
                           host);
    }
    else
        hipLaunchKernelGGL(k_zeroD, dim3(1), dim3(64), 0, ctx->stream, part + grid, int64_t(1));
    if (s->runCounter % s->updateInterval == 0)
        hipLaunchKernelGGL(k_updateMeanComp, dim3(gridFor(nh)), dim3(BLOCK), 0, ctx->stream, s->hist.p, s->hist.p + nh,
                           s->hist.p + 2 * nh, nh, 10.0);
    checkLaunch();
    s->igPosFresh = false;
    ++s->runCounter;
    if (energy)
    {
        // written by k_sumOrdered (0.0 from the host when there are no rows)
        sync(ctx);
        *energy = ctx->pinned[0];
    }
    MRMD_API_END
}

int mrmd_ljig_get_histogram(const mrmd_lj_idealgas* s, int which, double* host)
{
    auto* ss = const_cast<mrmd_lj_idealgas*>(s);
    MRMD_API_BEGIN(ss ? ss->ctx : nullptr)
    REQUIRE(ss && host && which >= 0 && which <= 2, "invalid arguments");
    const int64_t nh = COMP_BINS * ss->numTypes;
    HIPCHECK(hipMemcpyAsync(host, ss->hist.p + which * nh, size_t(nh) * 8, hipMemcpyDeviceToHost, ss->ctx->stream));
    sync(ss->ctx);
    MRMD_API_END
}

int mrmd_ljig_run_counter(mrmd_lj_idealgas* s, int64_t* get, const int64_t* set)
{
    if (!s || (!get && !set) || (set && *set < 0)) return MRMD_ERR_ARG;
    if (get) *get = s->runCounter;
    if (set) s->runCounter = *set;
    return MRMD_OK;
}

int mrmd_ljig_set_histogram(mrmd_lj_idealgas* s, int which, const double* host)
{
    MRMD_API_BEGIN(s ? s->ctx : nullptr)
    REQUIRE(s && host && which >= 0 && which <= 2, "invalid arguments");
    const int64_t nh = COMP_BINS * s->numTypes;
    HIPCHECK(hipMemcpyAsync(s->hist.p + which * nh, host, size_t(nh) * 8, hipMemcpyHostToDevice, s->ctx->stream));
    sync(s->ctx);
    MRMD_API_END
}

// ---------------------------------------------------------------------------------------
int mrmd_tf_create(mrmd_ctx* c, const double* target, const double* modulation, int64_t numTypes, const mrmd_box* box,
                   double binWidth, int sym, int periodic, mrmd_thermo_force** out)
{
    MRMD_API_BEGIN(c)
    REQUIRE(out && target && modulation && box && numTypes > 0 && binWidth > 0.0, "invalid arguments");
    const double diam0 = box->max[0] - box->min[0];
    const int64_t nb = int64_t(std::ceil(diam0 / binWidth));
    REQUIRE(nb >= 1, "no bins");
    REQUIRE(nb * numTypes <= MAX_TF_CELLS, "at most %d bins x types", MAX_TF_CELLS);
    auto* t = new mrmd_thermo_force();
    t->ctx = c;
    t->hd = makeHist(box->min[0], box->max[0], nb, numTypes);
    t->max = box->max[0];
    // binVolume = diameter[1] * diameter[2] * binSize (ThermodynamicForce.cpp:37)
    t->binVolume = (box->max[1] - box->min[1]) * (box->max[2] - box->min[2]) * t->hd.binSize;
    t->sym = sym != 0;
    t->periodic = periodic != 0;
    for (int64_t i = 0; i < numTypes; ++i) t->forceFactor.push_back(modulation[i] / target[i]);
    try
    {
        REQUIRE(t->hd.binSize <= binWidth, "requested bin size is not achieved");
        t->force.reserve(c, nb * numTypes);
        t->density.reserve(c, nb * numTypes);
        t->factor.reserve(c, numTypes);
        hipLaunchKernelGGL(k_zeroD, dim3(gridFor(nb * numTypes)), dim3(BLOCK), 0, c->stream, t->force.p, nb * numTypes);
        hipLaunchKernelGGL(k_zeroD, dim3(gridFor(nb * numTypes)), dim3(BLOCK), 0, c->stream, t->density.p, nb * numTypes);
        HIPCHECK(hipMemcpyAsync(t->factor.p, t->forceFactor.data(), size_t(numTypes) * 8, hipMemcpyHostToDevice, c->stream));
        checkLaunch();
        sync(c);
    }
    catch (...)
    {
        delete t;
        throw;
    }
    *out = t;
    MRMD_API_END
}

int mrmd_tf_destroy(mrmd_thermo_force* t)
{
    if (!t) return MRMD_OK;
    (void)hipStreamSynchronize(t->ctx->stream);
    delete t;
    return MRMD_OK;
}

int mrmd_tf_num_bins(const mrmd_thermo_force* t, int64_t* nb)
{
    if (!t || !nb) return MRMD_ERR_ARG;
    *nb = t->hd.numBins;
    return MRMD_OK;
}

int mrmd_tf_samples(const mrmd_thermo_force* t, int64_t* s)
{
    if (!t || !s) return MRMD_ERR_ARG;
    *s = t->samples;
    return MRMD_OK;
}

int mrmd_tf_set_density(mrmd_thermo_force* t, const double* host, int64_t samples)
{
    MRMD_API_BEGIN(t ? t->ctx : nullptr)
    REQUIRE(t && host && samples >= 0, "invalid arguments");
    const int64_t n = t->hd.numBins * t->hd.numHist;
    HIPCHECK(hipMemcpyAsync(t->density.p, host, size_t(n) * 8, hipMemcpyHostToDevice, t->ctx->stream));
    sync(t->ctx);
    t->samples = samples;
    MRMD_API_END
}

int mrmd_tf_sample(mrmd_thermo_force* t, const mrmd_atoms* a)
{
    MRMD_API_BEGIN(t ? t->ctx : nullptr)
    REQUIRE(t && a, "invalid arguments");
    REQUIRE(a->maxType < t->hd.numHist, "atom type outside the density profile");
    if (a->nlocal > 0)
    {
        KTimer k(t->ctx, "tf_sample");
        // about 4096 records per workgroup: every workgroup ends with one global atomic per
        // non-empty bin, and those contend on the few cache lines of the profile
        const unsigned grid = unsigned(std::min<int64_t>(std::max<int64_t>((a->nlocal + 4095) / 4096, 1), 256));
        hipLaunchKernelGGL(k_tfSample, dim3(grid), dim3(BLOCK), 0, t->ctx->stream, a->pos.p, a->nlocal, t->hd,
                           t->density.p);
        checkLaunch();
    }
    ++t->samples;
    MRMD_API_END
}

int mrmd_tf_update(mrmd_thermo_force* t, double sigma, double range, const mrmd_pred1* pred)
{
    MRMD_API_BEGIN(t ? t->ctx : nullptr)
    REQUIRE(t, "null thermodynamic force");
    validatePred1(pred);
    if (t->samples <= 0) fail(MRMD_ERR_STATE, "update without density samples");
    const double norm = 1.0 / (t->binVolume * double(t->samples));
    const int64_t delta = int64_t(int(range * sigma * t->hd.invBinSize));
    hipLaunchKernelGGL(k_tfUpdate, dim3(1), dim3(BLOCK), 0, t->ctx->stream, t->density.p, t->force.p, t->hd, norm,
                       t->factor.p, sigma, delta, t->sym ? 1 : 0, t->periodic ? 1 : 0, pred ? *pred : mrmd_pred1{},
                       pred && pred->kind != MRMD_PRED_ALWAYS ? 1 : 0);
    checkLaunch();
    t->samples = 0;
    MRMD_API_END
}

int mrmd_tf_apply(const mrmd_thermo_force* tc, mrmd_atoms* a, const mrmd_pred1* pred)
{
    auto* t = const_cast<mrmd_thermo_force*>(tc);
    MRMD_API_BEGIN(t ? t->ctx : nullptr)
    REQUIRE(t && a, "invalid arguments");
    REQUIRE(a->maxType < t->hd.numHist, "atom type outside the thermodynamic-force table");
    validatePred1(pred);
    if (a->nlocal == 0) return MRMD_OK;
    KTimer k(t->ctx, "tf_apply");
    const mrmd_pred1 p = pred ? *pred : mrmd_pred1{};
    if (pred && pred->kind != MRMD_PRED_ALWAYS)
        hipLaunchKernelGGL(k_tfApply<true>, dim3(gridFor(a->nlocal)), dim3(BLOCK), 0, t->ctx->stream, a->view(), a->nlocal,
                           t->force.p, t->hd, p);
    else
        hipLaunchKernelGGL(k_tfApply<false>, dim3(gridFor(a->nlocal)), dim3(BLOCK), 0, t->ctx->stream, a->view(),
                           a->nlocal, t->force.p, t->hd, p);
    checkLaunch();
    MRMD_API_END
}

int mrmd_tf_apply_interpolated(const mrmd_thermo_force* tc, mrmd_atoms* a, const mrmd_pred1* pred)
{
    auto* t = const_cast<mrmd_thermo_force*>(tc);
    MRMD_API_BEGIN(t ? t->ctx : nullptr)
    REQUIRE(t && a, "invalid arguments");
    REQUIRE(a->maxType < t->hd.numHist, "atom type outside the thermodynamic-force table");
    validatePred1(pred);
    if (a->nlocal == 0) return MRMD_OK;
    const mrmd_pred1 p = pred ? *pred : mrmd_pred1{};
    if (pred && pred->kind != MRMD_PRED_ALWAYS)
        hipLaunchKernelGGL(k_tfApplyInterp<true>, dim3(gridFor(a->nlocal)), dim3(BLOCK), 0, t->ctx->stream, a->view(),
                           a->nlocal, t->force.p, t->hd, p);
    else
        hipLaunchKernelGGL(k_tfApplyInterp<false>, dim3(gridFor(a->nlocal)), dim3(BLOCK), 0, t->ctx->stream, a->view(),
                           a->nlocal, t->force.p, t->hd, p);
    checkLaunch();
    MRMD_API_END
}

int mrmd_tf_get(const mrmd_thermo_force* tc, int which, double* host)
{
    auto* t = const_cast<mrmd_thermo_force*>(tc);
    MRMD_API_BEGIN(t ? t->ctx : nullptr)
    REQUIRE(t && host && (which == 0 || which == 1), "invalid arguments");
    const int64_t n = t->hd.numBins * t->hd.numHist;
    HIPCHECK(hipMemcpyAsync(host, which == 0 ? t->force.p : t->density.p, size_t(n) * 8, hipMemcpyDeviceToHost,
                            t->ctx->stream));
    sync(t->ctx);
    MRMD_API_END
}

int mrmd_tf_set_force(mrmd_thermo_force* t, const double* host)
{
    MRMD_API_BEGIN(t ? t->ctx : nullptr)
    REQUIRE(t && host, "invalid arguments");
    const int64_t n = t->hd.numBins * t->hd.numHist;
    HIPCHECK(hipMemcpyAsync(t->force.p, host, size_t(n) * 8, hipMemcpyHostToDevice, t->ctx->stream));
    sync(t->ctx);
    MRMD_API_END
}

int mrmd_tf_mu(const mrmd_thermo_force* tc, double* left, double* right)
{
    auto* t = const_cast<mrmd_thermo_force*>(tc);
    MRMD_API_BEGIN(t ? t->ctx : nullptr)
    REQUIRE(t && left && right, "invalid arguments");
    const int64_t nb = t->hd.numBins, nh = t->hd.numHist;
    std::vector<double> f(static_cast<size_t>(nb * nh));
    HIPCHECK(hipMemcpyAsync(f.data(), t->force.p, f.size() * 8, hipMemcpyDeviceToHost, t->ctx->stream));
    sync(t->ctx);
    for (int64_t h = 0; h < nh; ++h)
    {
        double l = 0.0, r = 0.0;
        for (int64_t b = 0; b < nb / 2; ++b) l += f[size_t(b * nh + h)];
        for (int64_t b = nb / 2; b < nb; ++b) r += f[size_t(b * nh + h)];
        left[h] = l * t->hd.binSize;
        right[h] = r * t->hd.binSize;
    }
    MRMD_API_END
}

int mrmd_hist_gradient(mrmd_ctx* c, double mn, double mx, int64_t nb, int64_t nh, const double* in, int periodic,
                       double* out)
{
    MRMD_API_BEGIN(c)
    REQUIRE(in && out && nb >= 2 && nh >= 1, "invalid arguments");
    DBuf<double> di, dout;
    di.reserve(c, nb * nh);
    dout.reserve(c, nb * nh);
    HIPCHECK(hipMemcpyAsync(di.p, in, size_t(nb * nh) * 8, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_histGradient, dim3(1), dim3(BLOCK), 0, c->stream, di.p, dout.p, makeHist(mn, mx, nb, nh),
                       periodic);
    checkLaunch();
    HIPCHECK(hipMemcpyAsync(out, dout.p, size_t(nb * nh) * 8, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    MRMD_API_END
}

int mrmd_hist_smoothen(mrmd_ctx* c, double mn, double mx, int64_t nb, int64_t nh, const double* in, double sigma,
                       double range, int periodic, double* out)
{
    MRMD_API_BEGIN(c)
    REQUIRE(in && out && nb >= 1 && nh >= 1 && sigma > 0.0, "invalid arguments");
    const HistD hd = makeHist(mn, mx, nb, nh);
    DBuf<double> di, dout;
    di.reserve(c, nb * nh);
    dout.reserve(c, nb * nh);
    HIPCHECK(hipMemcpyAsync(di.p, in, size_t(nb * nh) * 8, hipMemcpyHostToDevice, c->stream));
    const int64_t delta = int64_t(int(range * sigma * hd.invBinSize));
    hipLaunchKernelGGL(k_histSmoothen, dim3(1), dim3(BLOCK), 0, c->stream, di.p, dout.p, hd, sigma, delta, periodic);
    checkLaunch();
    HIPCHECK(hipMemcpyAsync(out, dout.p, size_t(nb * nh) * 8, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    MRMD_API_END
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// Fused passes of the AdResS time loop for molecules of one atom each, molecule k = atom k
// (createMoleculeForEachAtom; checked at every rebuild).  They compute exactly what the
// separate actions compute, in the same expressions:
//  * adressPrepare: UpdateGhostAtoms on the ghost records (non-rebuild steps),
//    UpdateMolecules with the slab weighting, the zero atom and molecule forces, and the
//    sorted-slot records the tiled LJ_IdealGas kernel reads;
//  * adressCollectGhosts: ContributeMoleculeForceToAtoms and AccumulateForce (fold) for the
//    ghost records (the local atoms' molecule forces are added in the fused post/pre step,
//    or by adressContributeLocal on a loop's last step).
namespace
{
__global__ void k_zeroOne(int32_t* p)
{
    if (threadIdx.x == 0) *p = 0;
}

__global__ void k_notIdentity(const int64_t* offset, const int64_t* numAtoms, int64_t n, int32_t* out)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n && (offset[k] != k || numAtoms[k] != 1)) atomicAdd(out, 1);
}

struct PrepArgs
{
    MolView m;
    AtomsView a;
    int64_t nlocal, nall;
    const int32_t* corr;  ///< ghost update from the real atoms, or null (rebuild step)
    double diam[3];
    mrmd_slab slab;
    double4* igPos;        ///< sorted-slot records for the tiled kernel, or null
    const int32_t* slotOf;
    int64_t rowsEnd;       ///< molecules [0, rowsEnd) are rows of the list
};

__global__ void k_adressPrepare(PrepArgs g)
{
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= g.nall) return;
    double4 p = g.a.pos[r];
    if (g.corr && r >= g.nlocal)
    {
        // UpdateGhostAtoms.cpp:31-68 (the expressions of k_ghostUpdate)
        const double4 q = g.a.pos[g.corr[r]];
        const double dx[3] = {p.x - q.x, p.y - q.y, p.z - q.z};
        double x[3] = {q.x, q.y, q.z};
        for (int d = 0; d < 3; ++d)
            if (dx[d] > +(0.1 * g.diam[d])) x[d] += g.diam[d];
        for (int d = 0; d < 3; ++d)
            if (dx[d] < -(0.1 * g.diam[d])) x[d] -= g.diam[d];
        p.x = x[0];
        p.y = x[1];
        p.z = x[2];
        g.a.pos[r] = p;
    }
    // UpdateMolecules::update for one atom (k_updateMolecules): the centre is relMass x the
    // atom position (summed from zero), then the slab weighting
    const double rm = g.a.relmass[r];
    double X = 0.0, Y = 0.0, Z = 0.0;
    X += p.x * rm;
    Y += p.y * rm;
    Z += p.z * rm;
    double4 q = g.m.pos[r];
    q.x = X;
    q.y = Y;
    q.z = Z;
    g.m.pos[r] = q;
    double l, md, gr;
    slabEval(g.slab, X, l, md, gr);
    g.m.lambda[r] = l;
    g.m.modLambda[r] = md;
    g.m.grad[r] = gr;
    g.m.grad[g.m.cap + r] = 0.0;
    g.m.grad[2 * g.m.cap + r] = 0.0;
    for (int d = 0; d < 3; ++d)
    {
        g.a.force[d * g.a.cap + r] = 0.0;
        g.m.force[d * g.m.cap + r] = 0.0;
    }
    // the record k_igGather writes (|lambda_mod|, negated for non-rows)
    if (g.igPos)
    {
        const double la = fabs(md);
        g.igPos[g.slotOf[r]] = make_double4(p.x, p.y, p.z, r < g.rowsEnd ? la : -la);
    }
}

// ghost records: molecule force to the ghost atom, then the fold into its real atom
// (k_contribute, then k_ghostFold: only non-zero components move)
__global__ void k_adressCollectGhosts(MolView m, AtomsView a, const int32_t* corr, int64_t begin, int64_t end)
{
    const int64_t i = begin + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= end) return;
    const double rm = a.relmass[i];
    double f[3];
    for (int d = 0; d < 3; ++d) f[d] = a.force[d * a.cap + i] + rm * m.force[d * m.cap + i];
    const int32_t real = corr[i];
    if (real == -1)
    {
        for (int d = 0; d < 3; ++d) a.force[d * a.cap + i] = f[d];
        return;
    }
    for (int d = 0; d < 3; ++d)
    {
        if (f[d] != 0.0) atomicAddD(&a.force[d * a.cap + real], f[d]);
        a.force[d * a.cap + i] = 0.0;
    }
}
}  // namespace

bool mrmd_hip::adressIdentityLayout(const mrmd_molecules* m, const mrmd_atoms* a)
{
    mrmd_ctx* ctx = m->ctx;
    const int64_t n = m->nlocal + m->nghost;
    if (n != a->nlocal + a->nghost || m->nlocal != a->nlocal) return false;
    if (n == 0) return true;
    int32_t* cnt = static_cast<int32_t*>(ctxScratch(ctx, 2, 64)) + 24;
    hipLaunchKernelGGL(k_zeroOne, dim3(1), dim3(64), 0, ctx->stream, cnt);
    hipLaunchKernelGGL(k_notIdentity, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, m->offset.p, m->numAtoms.p, n, cnt);
    checkLaunch();
    int32_t h[1];
    readbackI32(ctx, cnt, 1, h);
    return h[0] == 0;
}

void mrmd_hip::adressPrepare(mrmd_atoms* a, mrmd_molecules* m, const mrmd_ghost_layer* gl, const mrmd_box* box,
                             const mrmd_slab* slab, mrmd_lj_idealgas* ig, mrmd_verlet_list* vl)
{
    mrmd_ctx* ctx = a->ctx;
    const int64_t nall = a->nlocal + a->nghost;
    KTimer t(ctx, "adress_prepare");
    PrepArgs g{m->view(), a->view(), a->nlocal, nall, gl ? ghostCorrDevice(gl) : nullptr, {0, 0, 0}, *slab, nullptr,
               nullptr, std::min<int64_t>(m->nlocal, vl->end)};
    for (int d = 0; d < 3; ++d) g.diam[d] = box->max[d] - box->min[d];
    // the tiled LJ_IdealGas kernel's records, when that kernel will run on this list
    const bool tiled = ig->numTypes == 1 && vl->monoOf == m && vl->monoVersion == m->layoutVersion &&
                       vl->ntiles > 0 && vl->nrows == nall && !vl->globalStage;
    if (tiled)
    {
        ig->igPos.ensure(ctx, vl->nrows);
        g.igPos = ig->igPos.p;
        g.slotOf = vl->ensureSlotOf();
    }
    if (nall > 0) hipLaunchKernelGGL(k_adressPrepare, dim3(gridFor(nall)), dim3(BLOCK), 0, ctx->stream, g);
    checkLaunch();
    ig->igPosFresh = tiled;
    ig->igPosList = vl;
}

void mrmd_hip::adressCollectGhosts(mrmd_atoms* a, mrmd_molecules* m, const mrmd_ghost_layer* gl)
{
    mrmd_ctx* ctx = a->ctx;
    KTimer t(ctx, "adress_collect");
    if (a->nghost > 0)
        hipLaunchKernelGGL(k_adressCollectGhosts, dim3(gridFor(a->nghost)), dim3(BLOCK), 0, ctx->stream, m->view(),
                           a->view(), ghostCorrDevice(gl), a->nlocal, a->nlocal + a->nghost);
    checkLaunch();
}

void mrmd_hip::adressContributeLocal(mrmd_atoms* a, mrmd_molecules* m)
{
    mrmd_ctx* ctx = a->ctx;
    if (a->nlocal > 0)
        hipLaunchKernelGGL(k_contribute, dim3(gridFor(a->nlocal)), dim3(BLOCK), 0, ctx->stream, m->view(), a->view(),
                           a->nlocal);
    checkLaunch();
}

void mrmd_hip::thermoForceTable(const mrmd_thermo_force* tf, const double** force, HistD* hd)
{
    *force = tf->force.p;
    *hd = tf->hd;
}
