// Periodic halo on one device: wrap, ordered ghost creation, ghost refresh and force fold,
// for atoms (communication::GhostLayer) and molecules (communication::MultiResGhostLayer).
//
// Ghost creation keeps the reference's exact index order (GhostExchange.cpp:76-106, a
// Kokkos parallel_scan): per axis, the selected records of [0, nlocal + nghost) are
// appended in index order, first the block shifted by +L (x < minInner) then the block
// shifted by -L (x >= maxInner).  The compaction is a 1024-record tile count, a device
// scan over tiles, and a scatter pass ranking records inside the tile with wave ballots.
#include "containers.hpp"

using namespace mrmd_hip;

struct mrmd_ghost_layer
{
    mrmd_hip::CtxPtr ctx;
    DBuf<int32_t> corr;  ///< correspondingRealAtom, -1 for real atoms
    int64_t corrSize = 0;
    DBuf<int32_t> tileCnt;   ///< per-tile counts (2 arrays per axis, 26 for the three-axis pass)
    DBuf<int32_t> tileOff;   ///< their exclusive offsets
    DBuf<int32_t> blocks;    ///< three-axis pass: 26 totals, 27 bases by block code, the ghost count
};

namespace mrmd_hip
{
const int32_t* ghostCorrDevice(const mrmd_ghost_layer* gl) { return gl ? gl->corr.p : nullptr; }
int64_t ghostCorrSize(const mrmd_ghost_layer* gl) { return gl ? gl->corrSize : 0; }
}  // namespace mrmd_hip

namespace
{
constexpr int TILE = BLOCK * 4;

struct SubD
{
    double min[3], max[3], minInner[3], maxInner[3], diam[3];
};

// derived corners exactly as data::Subdomain (Subdomain.hpp:48-61)
SubD makeSub(const mrmd_box* b)
{
    SubD s;
    for (int d = 0; d < 3; ++d)
    {
        s.min[d] = b->min[d];
        s.max[d] = b->max[d];
        s.minInner[d] = b->min[d] + b->ghost[d];
        s.maxInner[d] = b->max[d] - b->ghost[d];
        s.diam[d] = b->max[d] - b->min[d];
    }
    return s;
}

// PeriodicMapping.cpp:30-58; from (if not null): the current positions are from[slotOf[i]]
__global__ void k_wrap(double4* pos, int64_t n, SubD s, const double4* from, const int32_t* slotOf)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 p = from ? from[slotOf[i]] : pos[i];
    wrapInto(p, s.min, s.max, s.diam);
    pos[i] = p;
}

__global__ void k_fillI32(int32_t* p, int64_t n, int32_t v)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// per tile: number of records with x < minInner (P, copied +L) and x >= maxInner (N, -L)
__global__ void k_ghostCount(const double4* pos, int64_t total, int axis, double lo, double hi,
                             int32_t* cntP, int32_t* cntN)
{
    __shared__ int sP[BLOCK / WAVE], sN[BLOCK / WAVE];
    const int64_t base = int64_t(blockIdx.x) * TILE;
    int cP = 0, cN = 0;
    for (int r = 0; r < 4; ++r)
    {
        const int64_t e = base + r * BLOCK + threadIdx.x;
        if (e < total)
        {
            const double x = coord(pos[e], axis);
            cP += x < lo;
            cN += x >= hi;
        }
    }
    for (int off = WAVE / 2; off > 0; off >>= 1)
    {
        cP += __shfl_xor(cP, off, WAVE);
        cN += __shfl_xor(cN, off, WAVE);
    }
    if (threadIdx.x % WAVE == 0)
    {
        sP[threadIdx.x / WAVE] = cP;
        sN[threadIdx.x / WAVE] = cN;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int a = 0, b = 0;
        for (int w = 0; w < BLOCK / WAVE; ++w)
        {
            a += sP[w];
            b += sN[w];
        }
        cntP[blockIdx.x] = a;
        cntN[blockIdx.x] = b;
    }
}

// GhostExchange.cpp:119-165: copy selected records in index order
__global__ void k_ghostScatter(AtomsView a, int32_t* corr, int64_t total, int axis, double lo, double hi,
                               double L, const int32_t* offP, const int32_t* offN, const int32_t* totP)
{
    __shared__ int wP[BLOCK / WAVE], wN[BLOCK / WAVE];
    const int64_t base = int64_t(blockIdx.x) * TILE;
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    const int64_t nP = *totP;
    int runP = 0, runN = 0;
    for (int r = 0; r < 4; ++r)
    {
        const int64_t e = base + r * BLOCK + threadIdx.x;
        bool fP = false, fN = false;
        if (e < total)
        {
            const double x = coord(a.pos[e], axis);
            fP = x < lo;
            fN = x >= hi;
        }
        const unsigned long long mP = __ballot(fP), mN = __ballot(fN);
        if (lane == 0)
        {
            wP[w] = __popcll(mP);
            wN[w] = __popcll(mN);
        }
        __syncthreads();
        int oP = 0, oN = 0, tP = 0, tN = 0;
        for (int i = 0; i < BLOCK / WAVE; ++i)
        {
            if (i < w)
            {
                oP += wP[i];
                oN += wN[i];
            }
            tP += wP[i];
            tN += wN[i];
        }
        __syncthreads();
        if (fP)
        {
            const int64_t g = total + offP[blockIdx.x] + runP + oP + waveRank(mP);
            copyAtom(a, g, e, axis, +L);
            const int32_t c = corr[e];
            corr[g] = c == -1 ? int32_t(e) : c;
        }
        if (fN)
        {
            const int64_t g = total + nP + offN[blockIdx.x] + runN + oN + waveRank(mN);
            copyAtom(a, g, e, axis, -L);
            const int32_t c = corr[e];
            corr[g] = c == -1 ? int32_t(e) : c;
        }
        runP += tP;
        runN += tN;
    }
}

// The three axes of GhostExchange::createGhostAtoms (GhostExchange.cpp:59-169) in one pass
// over the real atoms.  Running axis by axis, the exchange copies every record (real atoms,
// then the ghosts made so far, in index order) with coordinate < minInner by +L and >=
// maxInner by -L; a ghost keeps its source's other coordinates, so whether it is copied
// again along a later axis is its real atom's test along that axis.  Every ghost is thus
// one real atom i and a shift code b = sx + 3 sy + 9 sz (per axis 0: not copied, 1: +L,
// 2: -L; b != 0), and the exchange's output is 26 blocks, each the atoms of its code in
// index order, in the order ghostBlockOrder gives: x: +, -; y: + of (real, x+, x-), then
// - of the same; z: + of (real, x+, x-, y+ of the three, y- of the three), then - of the
// same.  Positions take each axis's shift once, as the axis-by-axis copies do.
// records per tile of the three-axis passes: rounds x BLOCK, up to four rounds as long as the
// grid keeps ~2,048 workgroups
inline int ghostRounds(int64_t n) { return int(std::min<int64_t>(4, std::max<int64_t>(1, n / (int64_t(BLOCK) * 2048)))); }
constexpr int GHOST_BLOCKS = 26;

__host__ __device__ inline int ghostBlockOrder(int b)
{
    const int sx = b % 3, sy = (b / 3) % 3, sz = b / 9;
    if (sz) return 8 + (sz - 1) * 9 + sy * 3 + sx;
    if (sy) return 2 + (sy - 1) * 3 + sx;
    return sx - 1;
}

// bit 2d: coordinate d < minInner (copied +L along d), bit 2d + 1: >= maxInner (-L)
__device__ inline unsigned ghostFlags(const double4& p, const SubD& s)
{
    return unsigned(p.x < s.minInner[0]) | (unsigned(p.x >= s.maxInner[0]) << 1) |
           (unsigned(p.y < s.minInner[1]) << 2) | (unsigned(p.y >= s.maxInner[1]) << 3) |
           (unsigned(p.z < s.minInner[2]) << 4) | (unsigned(p.z >= s.maxInner[2]) << 5);
}

// the codes an atom's flags put it in: bit b for every b != 0 with, per axis, s = 0 or the
// flag of s set
__device__ inline unsigned ghostCodes(unsigned f)
{
    const unsigned mx = 1u | ((f & 3u) << 1), my = 1u | (((f >> 2) & 3u) << 1), mz = 1u | (((f >> 4) & 3u) << 1);
    unsigned mxy = 0, m = 0;
    for (int k = 0; k < 3; ++k)
        if ((my >> k) & 1u) mxy |= mx << (3 * k);
    for (int k = 0; k < 3; ++k)
        if ((mz >> k) & 1u) m |= mxy << (9 * k);
    return f ? m & ~1u : 0u;
}

// the codes present in the wave
__device__ inline unsigned waveCodes(unsigned m)
{
    for (int off = WAVE / 2; off > 0; off >>= 1) m |= __shfl_xor(m, off, WAVE);
    return m;
}

// per tile of `rounds` x BLOCK atoms: atoms of each code, cnt[(b - 1) * tiles + tile]; a
// wave counts only the codes present in it
__global__ __launch_bounds__(BLOCK) void k_ghostCount3(const double4* pos, int64_t n, SubD s, int32_t* cnt,
                                                       int64_t tiles, int rounds)
{
    __shared__ int sc[BLOCK / WAVE][GHOST_BLOCKS];
    const int64_t base = int64_t(blockIdx.x) * rounds * BLOCK;
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    int acc = 0;  // lane b - 1: this wave's atoms of code b
    for (int r = 0; r < rounds; ++r)
    {
        const int64_t e = base + r * BLOCK + threadIdx.x;
        const unsigned codes = e < n ? ghostCodes(ghostFlags(pos[e], s)) : 0u;
        if (__ballot(codes != 0) == 0) continue;
        for (unsigned present = waveCodes(codes); present; present &= present - 1)
        {
            const int b = __builtin_ctz(present);
            const int c = __popcll(__ballot((codes >> b) & 1u));
            if (lane == b - 1) acc += c;
        }
    }
    if (lane < GHOST_BLOCKS) sc[w][lane] = acc;
    __syncthreads();
    if (threadIdx.x < GHOST_BLOCKS)
    {
        int t = 0;
        for (int q = 0; q < BLOCK / WAVE; ++q) t += sc[q][threadIdx.x];
        cnt[int64_t(threadIdx.x) * tiles + blockIdx.x] = t;
    }
}

// block bases in output order from the 26 totals (tot[b - 1]); bases[b] by code, *count = sum
__global__ void k_ghostBases(const int32_t* tot, int32_t* bases, int32_t* count)
{
    if (threadIdx.x != 0) return;
    int byOrder[GHOST_BLOCKS];
    for (int b = 1; b <= GHOST_BLOCKS; ++b) byOrder[ghostBlockOrder(b)] = tot[b - 1];
    int start[GHOST_BLOCKS], acc = 0;
    for (int o = 0; o < GHOST_BLOCKS; ++o)
    {
        start[o] = acc;
        acc += byOrder[o];
    }
    bases[0] = 0;
    for (int b = 1; b <= GHOST_BLOCKS; ++b) bases[b] = start[ghostBlockOrder(b)];
    *count = acc;
}

// every real atom writes its ghosts: ghost index n + bases[b] + the tile's offset + its rank
// in the tile among the atoms of code b (records of the tile in index order: round by round,
// wave by wave, lane by lane).  A flagged atom's record is read once, before its codes'
// copies; waves without a flagged atom only take part in the counts.
__global__ __launch_bounds__(BLOCK) void k_ghostScatter3(AtomsView a, int32_t* corr, int64_t n, SubD s,
                                                         const int32_t* off, int64_t tiles, const int32_t* bases,
                                                         int rounds)
{
    __shared__ int wc[BLOCK / WAVE][GHOST_BLOCKS];
    const int64_t base = int64_t(blockIdx.x) * rounds * BLOCK;
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    // lane b - 1: ghosts of code b of this tile written before this round
    int run = lane < GHOST_BLOCKS ? bases[lane + 1] + off[int64_t(lane) * tiles + blockIdx.x] : 0;
    for (int r = 0; r < rounds; ++r)
    {
        const int64_t e = base + r * BLOCK + threadIdx.x;
        double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
        unsigned codes = 0;
        if (e < n)
        {
            p = a.pos[e];
            codes = ghostCodes(ghostFlags(p, s));
        }
        double v[3] = {0.0, 0.0, 0.0}, f[3] = {0.0, 0.0, 0.0}, mass = 0.0, charge = 0.0, relmass = 0.0;
        if (codes)
        {
            for (int d = 0; d < 3; ++d)
            {
                v[d] = a.vel[d * a.cap + e];
                f[d] = a.force[d * a.cap + e];
            }
            mass = a.mass[e];
            charge = a.charge[e];
            relmass = a.relmass[e];
        }
        const unsigned present = __ballot(codes != 0) ? waveCodes(codes) : 0u;
        int mine = 0;
        for (unsigned pr = present; pr; pr &= pr - 1)
        {
            const int b = __builtin_ctz(pr);
            const int c = __popcll(__ballot((codes >> b) & 1u));
            if (lane == b - 1) mine = c;
        }
        if (lane < GHOST_BLOCKS) wc[w][lane] = mine;
        __syncthreads();
        int before = 0, total = 0;
        if (lane < GHOST_BLOCKS)
            for (int q = 0; q < BLOCK / WAVE; ++q)
            {
                const int c = wc[q][lane];
                before += q < w ? c : 0;
                total += c;
            }
        __syncthreads();
        for (unsigned pr = present; pr; pr &= pr - 1)
        {
            const int b = __builtin_ctz(pr);
            const bool m = (codes >> b) & 1u;
            const unsigned long long mask = __ballot(m);
            const int st = __shfl(run + before, b - 1, WAVE);
            if (m)
            {
                const int64_t g = n + st + waveRank(mask);
                const int sx = b % 3, sy = (b / 3) % 3, sz = b / 9;
                double4 q = p;
                if (sx) q.x += sx == 1 ? s.diam[0] : -s.diam[0];
                if (sy) q.y += sy == 1 ? s.diam[1] : -s.diam[1];
                if (sz) q.z += sz == 1 ? s.diam[2] : -s.diam[2];
                a.pos[g] = q;
                for (int d = 0; d < 3; ++d)
                {
                    a.vel[d * a.cap + g] = v[d];
                    a.force[d * a.cap + g] = f[d];
                }
                a.mass[g] = mass;
                a.charge[g] = charge;
                a.relmass[g] = relmass;
                corr[g] = int32_t(e);
            }
        }
        run += total;
    }
}

// UpdateGhostAtoms.cpp:31-68; sortedPos (if not null): the new position is also written to
// the record's sorted slot slotOf[i] of a Verlet list (the time loop's gather-free path)
// SLOTS: every record's current position is at its slot of sortedPos, and the ghost's new
// one is stored there alone
template <bool SLOTS>
__global__ void k_ghostUpdate(double4* pos, const int32_t* corr, int64_t begin, int64_t end, SubD s,
                              double4* sortedPos, const int32_t* slotOf)
{
    const int64_t i = begin + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= end) return;
    const int32_t real = corr[i];
    const int32_t slot = SLOTS ? slotOf[i] : 0;
    double4 p = SLOTS ? sortedPos[slot] : pos[i];
    const double4 q = SLOTS ? sortedPos[slotOf[real]] : pos[real];
    const double dx[3] = {p.x - q.x, p.y - q.y, p.z - q.z};
    double x[3] = {q.x, q.y, q.z};
    for (int d = 0; d < 3; ++d)
        if (dx[d] > +(0.1 * s.diam[d])) x[d] += s.diam[d];
    for (int d = 0; d < 3; ++d)
        if (dx[d] < -(0.1 * s.diam[d])) x[d] -= s.diam[d];
    p.x = x[0];
    p.y = x[1];
    p.z = x[2];
    if (SLOTS)
    {
        sortedPos[slot] = p;
        return;
    }
    pos[i] = p;
    if (sortedPos) sortedPos[slotOf[i]] = p;
}

// AccumulateForce.cpp:25-47
__global__ void k_ghostFold(double* force, int64_t cap, const int32_t* corr, int64_t begin, int64_t end)
{
    const int64_t i = begin + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= end) return;
    const int32_t real = corr[i];
    if (real == -1) return;
    for (int d = 0; d < 3; ++d)
    {
        const double f = force[d * cap + i];
        if (f != 0.0)
        {
            atomicAddD(&force[d * cap + real], f);
            force[d * cap + i] = 0.0;
        }
    }
}

// MultiResRealAtomsExchange.cpp:23-73
__global__ void k_molWrap(MolView m, AtomsView a, int64_t n, SubD s)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 p = m.pos[i];
    double X[3] = {p.x, p.y, p.z};
    const int64_t start = m.offset[i], end = start + m.numAtoms[i];
    for (int d = 0; d < 3; ++d)
    {
        double shift = 0.0;
        if (s.max[d] <= X[d])
        {
            X[d] -= s.diam[d];
            shift = -s.diam[d];
            for (int64_t k = start; k < end; ++k)
            {
                double4 q = a.pos[k];
                if (d == 0) q.x -= s.diam[d];
                if (d == 1) q.y -= s.diam[d];
                if (d == 2) q.z -= s.diam[d];
                a.pos[k] = q;
            }
        }
        if (X[d] < s.min[d])
        {
            X[d] += s.diam[d];
            for (int64_t k = start; k < end; ++k)
            {
                double4 q = a.pos[k];
                if (d == 0) q.x += s.diam[d];
                if (d == 1) q.y += s.diam[d];
                if (d == 2) q.z += s.diam[d];
                a.pos[k] = q;
            }
        }
        (void)shift;
    }
    p.x = X[0];
    p.y = X[1];
    p.z = X[2];
    m.pos[i] = p;
}

// per tile: molecules / atoms with centre >= maxInner (P) and < minInner (N)
__global__ void k_molGhostCount(MolView m, int64_t total, int axis, double lo, double hi, int32_t* cnt,
                                int64_t tiles)
{
    __shared__ int sh[4][BLOCK / WAVE];
    const int64_t base = int64_t(blockIdx.x) * TILE;
    int c[4] = {0, 0, 0, 0};  // molP, atomP, molN, atomN
    for (int r = 0; r < 4; ++r)
    {
        const int64_t e = base + r * BLOCK + threadIdx.x;
        if (e < total)
        {
            const double x = coord(m.pos[e], axis);
            const int na = int(m.numAtoms[e]);
            if (x >= hi)
            {
                c[0] += 1;
                c[1] += na;
            }
            if (x < lo)
            {
                c[2] += 1;
                c[3] += na;
            }
        }
    }
    for (int k = 0; k < 4; ++k)
    {
        int v = c[k];
        for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
        if (threadIdx.x % WAVE == 0) sh[k][threadIdx.x / WAVE] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4)
    {
        int v = 0;
        for (int w = 0; w < BLOCK / WAVE; ++w) v += sh[threadIdx.x][w];
        cnt[threadIdx.x * tiles + blockIdx.x] = v;
    }
}

__device__ inline void copyMol(const MolView& m, int64_t dst, int64_t src, int axis, double shift)
{
    double4 p = m.pos[src];
    if (axis == 0) p.x += shift;
    if (axis == 1) p.y += shift;
    if (axis == 2) p.z += shift;
    m.pos[dst] = p;
    for (int d = 0; d < 3; ++d)
    {
        m.force[d * m.cap + dst] = m.force[d * m.cap + src];
        m.grad[d * m.cap + dst] = m.grad[d * m.cap + src];
    }
    m.lambda[dst] = m.lambda[src];
    m.modLambda[dst] = m.modLambda[src];
}

// MultiResPeriodicGhostExchange.cpp:140-240
__global__ void k_molGhostScatter(MolView m, AtomsView a, int32_t* corr, int64_t totM, int64_t totA, int axis,
                                  double lo, double hi, double L, const int32_t* off, int64_t tiles,
                                  const int32_t* tot)
{
    __shared__ int lds[BLOCK / WAVE];
    const int64_t base = int64_t(blockIdx.x) * TILE;
    const int64_t nPM = tot[0], nPA = tot[1];
    int run[4] = {0, 0, 0, 0};
    for (int r = 0; r < 4; ++r)
    {
        const int64_t e = base + r * BLOCK + threadIdx.x;
        bool fP = false, fN = false;
        int na = 0;
        if (e < totM)
        {
            const double x = coord(m.pos[e], axis);
            fP = x >= hi;
            fN = x < lo;
            na = int(m.numAtoms[e]);
        }
        int t[4];
        const int ex0 = blockExclusiveScan<int>(fP ? 1 : 0, lds, &t[0]);
        const int ex1 = blockExclusiveScan<int>(fP ? na : 0, lds, &t[1]);
        const int ex2 = blockExclusiveScan<int>(fN ? 1 : 0, lds, &t[2]);
        const int ex3 = blockExclusiveScan<int>(fN ? na : 0, lds, &t[3]);
        if (fP || fN)
        {
            const int64_t srcOff = m.offset[e];
            int64_t dstM, dstA;
            double shift;
            if (fP)
            {
                dstM = totM + off[0 * tiles + blockIdx.x] + run[0] + ex0;
                dstA = totA + off[1 * tiles + blockIdx.x] + run[1] + ex1;
                shift = -L;
            }
            else
            {
                dstM = totM + nPM + off[2 * tiles + blockIdx.x] + run[2] + ex2;
                dstA = totA + nPA + off[3 * tiles + blockIdx.x] + run[3] + ex3;
                shift = +L;
            }
            // a molecule can satisfy both only if the box is thinner than two ghost layers;
            // the reference then copies it into both blocks
            for (int pass = 0; pass < (fP && fN ? 2 : 1); ++pass)
            {
                if (pass == 1)
                {
                    dstM = totM + nPM + off[2 * tiles + blockIdx.x] + run[2] + ex2;
                    dstA = totA + nPA + off[3 * tiles + blockIdx.x] + run[3] + ex3;
                    shift = +L;
                }
                copyMol(m, dstM, e, axis, shift);
                m.offset[dstM] = dstA;
                m.numAtoms[dstM] = na;
                for (int k = 0; k < na; ++k)
                {
                    copyAtom(a, dstA + k, srcOff + k, axis, shift);
                    const int32_t c = corr[srcOff + k];
                    corr[dstA + k] = c == -1 ? int32_t(srcOff + k) : c;
                }
            }
        }
        for (int k = 0; k < 4; ++k) run[k] += t[k];
    }
}
// The multi-resolution exchange (MultiResPeriodicGhostExchange.cpp:140-240) of the three
// axes in one pass over the real molecules, as k_ghostCount3 / k_ghostScatter3 do for
// atoms: per axis the reference copies the molecules with centre >= maxInner by -L (its
// first block) and those < minInner by +L, each with its atoms in a run after the atoms made
// so far; a ghost molecule's later tests are its real molecule's, so the output is the 26
// code blocks in ghostBlockOrder (code s = 1 here: >= maxInner, -L; 2: < minInner, +L),
// molecules and their atoms block by block, both in real-molecule order.
__device__ inline unsigned molGhostFlags(const double4& p, const SubD& s)
{
    return unsigned(p.x >= s.maxInner[0]) | (unsigned(p.x < s.minInner[0]) << 1) |
           (unsigned(p.y >= s.maxInner[1]) << 2) | (unsigned(p.y < s.minInner[1]) << 3) |
           (unsigned(p.z >= s.maxInner[2]) << 4) | (unsigned(p.z < s.minInner[2]) << 5);
}

// per tile of `rounds` x BLOCK molecules: molecules of each code, cnt[(b - 1) * tiles + tile], and their
// atoms, cnt[(26 + b - 1) * tiles + tile]
__global__ __launch_bounds__(BLOCK) void k_molGhostCount3(MolView m, int64_t n, SubD s, int32_t* cnt,
                                                          int64_t tiles, int rounds)
{
    __shared__ int sc[BLOCK / WAVE][2 * GHOST_BLOCKS];
    const int64_t base = int64_t(blockIdx.x) * rounds * BLOCK;
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    int accM = 0, accA = 0;  // lane b - 1: this wave's molecules / atoms of code b
    for (int r = 0; r < rounds; ++r)
    {
        const int64_t e = base + r * BLOCK + threadIdx.x;
        unsigned codes = 0;
        int na = 0;
        if (e < n)
        {
            codes = ghostCodes(molGhostFlags(m.pos[e], s));
            na = int(m.numAtoms[e]);
        }
        if (__ballot(codes != 0) == 0) continue;
        for (unsigned present = waveCodes(codes); present; present &= present - 1)
        {
            const int b = __builtin_ctz(present);
            const bool in = (codes >> b) & 1u;
            const int cm = __popcll(__ballot(in));
            int ca = in ? na : 0;
            for (int off = WAVE / 2; off > 0; off >>= 1) ca += __shfl_xor(ca, off, WAVE);
            if (lane == b - 1)
            {
                accM += cm;
                accA += ca;
            }
        }
    }
    if (lane < GHOST_BLOCKS)
    {
        sc[w][lane] = accM;
        sc[w][GHOST_BLOCKS + lane] = accA;
    }
    __syncthreads();
    if (threadIdx.x < 2 * GHOST_BLOCKS)
    {
        int t = 0;
        for (int q = 0; q < BLOCK / WAVE; ++q) t += sc[q][threadIdx.x];
        cnt[int64_t(threadIdx.x) * tiles + blockIdx.x] = t;
    }
}

// molecule and atom block bases in output order (tot[b - 1] molecules, tot[26 + b - 1] atoms);
// bases[b] and bases[27 + b] by code, count[0] molecules, count[1] atoms
__global__ void k_molGhostBases(const int32_t* tot, int32_t* bases, int32_t* count)
{
    if (threadIdx.x >= 2) return;
    const int k = threadIdx.x;  // 0: molecules, 1: atoms
    int byOrder[GHOST_BLOCKS];
    for (int b = 1; b <= GHOST_BLOCKS; ++b) byOrder[ghostBlockOrder(b)] = tot[k * GHOST_BLOCKS + b - 1];
    int start[GHOST_BLOCKS], acc = 0;
    for (int o = 0; o < GHOST_BLOCKS; ++o)
    {
        start[o] = acc;
        acc += byOrder[o];
    }
    bases[k * 27] = 0;
    for (int b = 1; b <= GHOST_BLOCKS; ++b) bases[k * 27 + b] = start[ghostBlockOrder(b)];
    count[k] = acc;
}

// every real molecule writes its ghost molecules and their atoms (molecule record and its
// first atom read once, before its codes' copies; further atoms of a molecule per copy)
__global__ __launch_bounds__(BLOCK) void k_molGhostScatter3(MolView m, AtomsView a, int32_t* corr, int64_t nM,
                                                            int64_t nA, SubD s, const int32_t* off, int64_t tiles,
                                                            const int32_t* bases, int rounds)
{
    __shared__ int wc[BLOCK / WAVE][2 * GHOST_BLOCKS];
    const int64_t base = int64_t(blockIdx.x) * rounds * BLOCK;
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    // lane b - 1: molecules / atoms of code b of this tile written before this round
    int runM = 0, runA = 0;
    if (lane < GHOST_BLOCKS)
    {
        runM = bases[lane + 1] + off[int64_t(lane) * tiles + blockIdx.x];
        runA = bases[27 + lane + 1] + off[int64_t(GHOST_BLOCKS + lane) * tiles + blockIdx.x];
    }
    for (int r = 0; r < rounds; ++r)
    {
        const int64_t e = base + r * BLOCK + threadIdx.x;
        double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
        unsigned codes = 0;
        int na = 0;
        if (e < nM)
        {
            p = m.pos[e];
            codes = ghostCodes(molGhostFlags(p, s));
            na = int(m.numAtoms[e]);
        }
        // the molecule's record and its first atom's, for all its copies
        double mf[3] = {0.0, 0.0, 0.0}, mg[3] = {0.0, 0.0, 0.0}, lam = 0.0, mod = 0.0;
        int64_t src = 0;
        double4 ap = make_double4(0.0, 0.0, 0.0, 0.0);
        double av[3] = {0.0, 0.0, 0.0}, af[3] = {0.0, 0.0, 0.0}, amass = 0.0, acharge = 0.0, arel = 0.0;
        if (codes)
        {
            for (int d = 0; d < 3; ++d)
            {
                mf[d] = m.force[d * m.cap + e];
                mg[d] = m.grad[d * m.cap + e];
            }
            lam = m.lambda[e];
            mod = m.modLambda[e];
            src = m.offset[e];
            if (na > 0)
            {
                ap = a.pos[src];
                for (int d = 0; d < 3; ++d)
                {
                    av[d] = a.vel[d * a.cap + src];
                    af[d] = a.force[d * a.cap + src];
                }
                amass = a.mass[src];
                acharge = a.charge[src];
                arel = a.relmass[src];
            }
        }
        const unsigned present = __ballot(codes != 0) ? waveCodes(codes) : 0u;
        int mineM = 0, mineA = 0;
        for (unsigned pr = present; pr; pr &= pr - 1)
        {
            const int b = __builtin_ctz(pr);
            const bool in = (codes >> b) & 1u;
            const int cm = __popcll(__ballot(in));
            int ca = in ? na : 0;
            for (int o = WAVE / 2; o > 0; o >>= 1) ca += __shfl_xor(ca, o, WAVE);
            if (lane == b - 1)
            {
                mineM = cm;
                mineA = ca;
            }
        }
        if (lane < GHOST_BLOCKS)
        {
            wc[w][lane] = mineM;
            wc[w][GHOST_BLOCKS + lane] = mineA;
        }
        __syncthreads();
        int beforeM = 0, beforeA = 0, totalM = 0, totalA = 0;
        if (lane < GHOST_BLOCKS)
            for (int q = 0; q < BLOCK / WAVE; ++q)
            {
                const int cm = wc[q][lane], ca = wc[q][GHOST_BLOCKS + lane];
                beforeM += q < w ? cm : 0;
                beforeA += q < w ? ca : 0;
                totalM += cm;
                totalA += ca;
            }
        __syncthreads();
        for (unsigned pr = present; pr; pr &= pr - 1)
        {
            const int b = __builtin_ctz(pr);
            const bool in = (codes >> b) & 1u;
            const unsigned long long mask = __ballot(in);
            const int stM = __shfl(runM + beforeM, b - 1, WAVE), stA = __shfl(runA + beforeA, b - 1, WAVE);
            const int inc = waveInclusiveScan(in ? na : 0);
            if (in)
            {
                const int64_t dM = nM + stM + waveRank(mask);
                const int64_t dA = nA + stA + (inc - na);
                const int sx = b % 3, sy = (b / 3) % 3, sz = b / 9;
                // code 1: copied by -L, code 2: by +L
                const double ox = sx ? (sx == 1 ? -s.diam[0] : s.diam[0]) : 0.0;
                const double oy = sy ? (sy == 1 ? -s.diam[1] : s.diam[1]) : 0.0;
                const double oz = sz ? (sz == 1 ? -s.diam[2] : s.diam[2]) : 0.0;
                double4 q = p;
                if (sx) q.x += ox;
                if (sy) q.y += oy;
                if (sz) q.z += oz;
                m.pos[dM] = q;
                for (int d = 0; d < 3; ++d)
                {
                    m.force[d * m.cap + dM] = mf[d];
                    m.grad[d * m.cap + dM] = mg[d];
                }
                m.lambda[dM] = lam;
                m.modLambda[dM] = mod;
                m.offset[dM] = dA;
                m.numAtoms[dM] = na;
                for (int k = 0; k < na; ++k)
                {
                    double4 t = ap;
                    double tv[3] = {av[0], av[1], av[2]}, tf[3] = {af[0], af[1], af[2]};
                    double tm = amass, tc = acharge, tr = arel;
                    if (k > 0)
                    {
                        t = a.pos[src + k];
                        for (int d = 0; d < 3; ++d)
                        {
                            tv[d] = a.vel[d * a.cap + src + k];
                            tf[d] = a.force[d * a.cap + src + k];
                        }
                        tm = a.mass[src + k];
                        tc = a.charge[src + k];
                        tr = a.relmass[src + k];
                    }
                    if (sx) t.x += ox;
                    if (sy) t.y += oy;
                    if (sz) t.z += oz;
                    a.pos[dA + k] = t;
                    for (int d = 0; d < 3; ++d)
                    {
                        a.vel[d * a.cap + dA + k] = tv[d];
                        a.force[d * a.cap + dA + k] = tf[d];
                    }
                    a.mass[dA + k] = tm;
                    a.charge[dA + k] = tc;
                    a.relmass[dA + k] = tr;
                    corr[dA + k] = int32_t(src + k);
                }
            }
        }
        runM += totalM;
        runA += totalA;
    }
}
}  // namespace

// =========================================================================================
namespace
{
void resetCorr(mrmd_ghost_layer* gl, int64_t nlocal)
{
    gl->corr.reserve(gl->ctx, growCap(nlocal), 0);
    if (nlocal > 0)
        hipLaunchKernelGGL(k_fillI32, dim3(gridFor(nlocal)), dim3(BLOCK), 0, gl->ctx->stream, gl->corr.p, nlocal,
                           -1);
    gl->corrSize = nlocal;
    checkLaunch();
}

void createAxis(mrmd_ghost_layer* gl, mrmd_atoms* a, const mrmd_box* box, int axis)
{
    mrmd_ctx* ctx = gl->ctx;
    const SubD s = makeSub(box);
    const int64_t total = a->nlocal + a->nghost;
    if (total >= (int64_t(1) << 31) - 1) fail(MRMD_ERR_UNSUP, "more than 2^31 records");
    const int64_t tiles = std::max<int64_t>(1, (total + TILE - 1) / TILE);
    gl->tileCnt.ensure(ctx, 2 * tiles);
    gl->tileOff.ensure(ctx, 2 * tiles);
    int32_t* tot = static_cast<int32_t*>(ctxScratch(ctx, 2, 64));
    hipLaunchKernelGGL(k_ghostCount, dim3(unsigned(tiles)), dim3(BLOCK), 0, ctx->stream, a->pos.p, total, axis,
                       s.minInner[axis], s.maxInner[axis], gl->tileCnt.p, gl->tileCnt.p + tiles);
    checkLaunch();
    scanExclusiveI32Multi(ctx, gl->tileCnt.p, gl->tileOff.p, tiles, 2, tot);
    int32_t h[2];
    readbackI32(ctx, tot, 2, h);
    const int64_t nP = h[0], nN = h[1];
    const int64_t newTotal = total + nP + nN;
    // atoms.resize(...) keeps existing records (GhostExchange.cpp:111-112)
    a->ensureCap(newTotal);
    a->size = newTotal;
    if (gl->corr.n < newTotal) gl->corr.reserve(ctx, growCap(newTotal), gl->corrSize);
    if (nP + nN > 0)
    {
        hipLaunchKernelGGL(k_ghostScatter, dim3(unsigned(tiles)), dim3(BLOCK), 0, ctx->stream, a->view(), gl->corr.p,
                           total, axis, s.minInner[axis], s.maxInner[axis], s.diam[axis], gl->tileOff.p,
                           gl->tileOff.p + tiles, tot);
        checkLaunch();
    }
    gl->corrSize = newTotal;
    a->nghost += nP + nN;
}

// the three axes in one pass over the real atoms (no ghosts yet): the same ghosts in the same
// order as createAxis for x, y and z
void createAll(mrmd_ghost_layer* gl, mrmd_atoms* a, const mrmd_box* box)
{
    mrmd_ctx* ctx = gl->ctx;
    const SubD s = makeSub(box);
    const int64_t n = a->nlocal;
    if (n >= (int64_t(1) << 31) - 1) fail(MRMD_ERR_UNSUP, "more than 2^31 records");
    const int rounds = ghostRounds(n);
    const int64_t tiles = std::max<int64_t>(1, (n + rounds * BLOCK - 1) / (rounds * BLOCK));
    gl->tileCnt.ensure(ctx, GHOST_BLOCKS * tiles);
    gl->tileOff.ensure(ctx, GHOST_BLOCKS * tiles);
    gl->blocks.ensure(ctx, 64);
    int32_t* tot = gl->blocks.p;
    int32_t* bases = tot + 32;
    int32_t* count = bases + 27;
    hipLaunchKernelGGL(k_ghostCount3, dim3(unsigned(tiles)), dim3(BLOCK), 0, ctx->stream, a->pos.p, n, s,
                       gl->tileCnt.p, tiles, rounds);
    checkLaunch();
    scanExclusiveI32Multi(ctx, gl->tileCnt.p, gl->tileOff.p, tiles, GHOST_BLOCKS, tot);
    hipLaunchKernelGGL(k_ghostBases, dim3(1), dim3(64), 0, ctx->stream, tot, bases, count);
    checkLaunch();
    int32_t h[1];
    readbackI32(ctx, count, 1, h);
    const int64_t newTotal = n + h[0];
    a->ensureCap(newTotal);
    a->size = newTotal;
    if (gl->corr.n < newTotal) gl->corr.reserve(ctx, growCap(newTotal), gl->corrSize);
    if (h[0] > 0)
    {
        hipLaunchKernelGGL(k_ghostScatter3, dim3(unsigned(tiles)), dim3(BLOCK), 0, ctx->stream, a->view(), gl->corr.p,
                           n, s, gl->tileOff.p, tiles, bases, rounds);
        checkLaunch();
    }
    gl->corrSize = newTotal;
    a->nghost = h[0];
}

void molCreateAxis(mrmd_ghost_layer* gl, mrmd_molecules* m, mrmd_atoms* a, const mrmd_box* box, int axis)
{
    mrmd_ctx* ctx = gl->ctx;
    const SubD s = makeSub(box);
    const int64_t totM = m->nlocal + m->nghost;
    const int64_t totA = a->nlocal + a->nghost;
    const int64_t tiles = std::max<int64_t>(1, (totM + TILE - 1) / TILE);
    gl->tileCnt.ensure(ctx, 4 * tiles);
    gl->tileOff.ensure(ctx, 4 * tiles);
    int32_t* tot = static_cast<int32_t*>(ctxScratch(ctx, 2, 64));
    hipLaunchKernelGGL(k_molGhostCount, dim3(unsigned(tiles)), dim3(BLOCK), 0, ctx->stream, m->view(), totM, axis,
                       s.minInner[axis], s.maxInner[axis], gl->tileCnt.p, tiles);
    checkLaunch();
    scanExclusiveI32Multi(ctx, gl->tileCnt.p, gl->tileOff.p, tiles, 4, tot);
    int32_t h[4];
    readbackI32(ctx, tot, 4, h);
    const int64_t newM = totM + h[0] + h[2], newA = totA + h[1] + h[3];
    m->ensureCap(newM);
    m->size = newM;
    a->ensureCap(newA);
    a->size = newA;
    if (gl->corr.n < newA) gl->corr.reserve(ctx, growCap(newA), gl->corrSize);
    if (h[0] + h[2] > 0)
    {
        hipLaunchKernelGGL(k_molGhostScatter, dim3(unsigned(tiles)), dim3(BLOCK), 0, ctx->stream, m->view(), a->view(),
                           gl->corr.p, totM, totA, axis, s.minInner[axis], s.maxInner[axis], s.diam[axis],
                           gl->tileOff.p, tiles, tot);
        checkLaunch();
    }
    gl->corrSize = newA;
    m->nghost += h[0] + h[2];
    ++m->layoutVersion;
    a->nghost += h[1] + h[3];
}

// the three axes in one pass over the real molecules (no ghosts yet): the same ghost
// molecules and atoms in the same order as molCreateAxis for x, y and z
void molCreateAll(mrmd_ghost_layer* gl, mrmd_molecules* m, mrmd_atoms* a, const mrmd_box* box)
{
    mrmd_ctx* ctx = gl->ctx;
    const SubD s = makeSub(box);
    const int64_t nM = m->nlocal, nA = a->nlocal;
    const int rounds = ghostRounds(nM);
    const int64_t tiles = std::max<int64_t>(1, (nM + rounds * BLOCK - 1) / (rounds * BLOCK));
    gl->tileCnt.ensure(ctx, 2 * GHOST_BLOCKS * tiles);
    gl->tileOff.ensure(ctx, 2 * GHOST_BLOCKS * tiles);
    gl->blocks.ensure(ctx, 128);
    int32_t* tot = gl->blocks.p;
    int32_t* bases = tot + 56;
    int32_t* count = bases + 56;
    hipLaunchKernelGGL(k_molGhostCount3, dim3(unsigned(tiles)), dim3(BLOCK), 0, ctx->stream, m->view(), nM, s,
                       gl->tileCnt.p, tiles, rounds);
    checkLaunch();
    scanExclusiveI32Multi(ctx, gl->tileCnt.p, gl->tileOff.p, tiles, 2 * GHOST_BLOCKS, tot);
    hipLaunchKernelGGL(k_molGhostBases, dim3(1), dim3(64), 0, ctx->stream, tot, bases, count);
    checkLaunch();
    int32_t h[2];
    readbackI32(ctx, count, 2, h);
    const int64_t newM = nM + h[0], newA = nA + h[1];
    m->ensureCap(newM);
    m->size = newM;
    a->ensureCap(newA);
    a->size = newA;
    if (gl->corr.n < newA) gl->corr.reserve(ctx, growCap(newA), gl->corrSize);
    if (h[0] > 0)
    {
        hipLaunchKernelGGL(k_molGhostScatter3, dim3(unsigned(tiles)), dim3(BLOCK), 0, ctx->stream, m->view(),
                           a->view(), gl->corr.p, nM, nA, s, gl->tileOff.p, tiles, bases, rounds);
        checkLaunch();
    }
    gl->corrSize = newA;
    m->nghost = h[0];
    ++m->layoutVersion;
    a->nghost = h[1];
}

void checkBox(const mrmd_box* box)
{
    REQUIRE(box, "null box");
    for (int d = 0; d < 3; ++d)
    {
        REQUIRE(box->max[d] > box->min[d], "box max must exceed min");
        REQUIRE(box->ghost[d] >= 0.0, "negative ghost layer");
        REQUIRE(box->max[d] - box->min[d] > box->ghost[d], "ghost layer larger than subdomain");
    }
}
}  // namespace

// UpdateGhostAtoms::updateOnlyPos with each new ghost position also written to its sorted
// slot of a Verlet list (time loops that keep the list's slot-order copy current)
int mrmd_hip::ghostUpdateSorted(mrmd_ghost_layer* gl, mrmd_atoms* a, const mrmd_box* box, double4* sortedPos,
                                const int32_t* slotOf, bool slotsOnly)
{
    MRMD_API_BEGIN(gl ? gl->ctx : nullptr)
    REQUIRE(gl && a && sortedPos && slotOf, "invalid arguments");
    checkBox(box);
    REQUIRE(gl->corrSize >= a->nlocal + a->nghost, "correspondingRealAtom shorter than the atoms");
    if (a->nghost == 0) return MRMD_OK;
    KTimer t(gl->ctx, "halo_update");
    auto kernel = slotsOnly ? k_ghostUpdate<true> : k_ghostUpdate<false>;
    hipLaunchKernelGGL(kernel, dim3(gridFor(a->nghost)), dim3(BLOCK), 0, gl->ctx->stream, a->pos.p, gl->corr.p,
                       a->nlocal, a->nlocal + a->nghost, makeSub(box), sortedPos, slotOf);
    checkLaunch();
    MRMD_API_END
}

// PeriodicMapping of the local atoms from their slots of a list's slot-order copy
int mrmd_hip::haloWrapFromSlots(mrmd_atoms* a, const mrmd_box* box, const double4* from, const int32_t* slotOf)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && from && slotOf, "invalid arguments");
    checkBox(box);
    if (a->nlocal == 0) return MRMD_OK;
    KTimer t(a->ctx, "halo_wrap");
    hipLaunchKernelGGL(k_wrap, dim3(gridFor(a->nlocal)), dim3(BLOCK), 0, a->ctx->stream, a->pos.p, a->nlocal,
                       makeSub(box), from, slotOf);
    checkLaunch();
    MRMD_API_END
}

extern "C" {

int mrmd_ghost_layer_create(mrmd_ctx* c, mrmd_ghost_layer** out)
{
    MRMD_API_BEGIN(c)
    REQUIRE(out, "null output");
    auto* g = new mrmd_ghost_layer();
    g->ctx = c;
    *out = g;
    MRMD_API_END
}

int mrmd_ghost_layer_destroy(mrmd_ghost_layer* g)
{
    if (!g) return MRMD_OK;
    (void)hipStreamSynchronize(g->ctx->stream);
    delete g;
    return MRMD_OK;
}

int mrmd_halo_wrap(mrmd_atoms* a, const mrmd_box* box)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a, "null atoms");
    checkBox(box);
    if (a->nlocal == 0) return MRMD_OK;
    KTimer t(a->ctx, "halo_wrap");
    hipLaunchKernelGGL(k_wrap, dim3(gridFor(a->nlocal)), dim3(BLOCK), 0, a->ctx->stream, a->pos.p, a->nlocal,
                       makeSub(box), nullptr, nullptr);
    checkLaunch();
    MRMD_API_END
}

int mrmd_halo_create(mrmd_ghost_layer* gl, mrmd_atoms* a, const mrmd_box* box, int64_t* nghost)
{
    MRMD_API_BEGIN(gl ? gl->ctx : nullptr)
    REQUIRE(gl && a && a->ctx == gl->ctx, "invalid arguments");
    checkBox(box);
    KTimer t(gl->ctx, "halo_create");
    resetCorr(gl, a->nlocal);
    a->nghost = 0;
    createAll(gl, a, box);
    if (nghost) *nghost = a->nghost;
    MRMD_API_END
}

int mrmd_halo_create_axis(mrmd_ghost_layer* gl, mrmd_atoms* a, const mrmd_box* box, int axis, int reset)
{
    MRMD_API_BEGIN(gl ? gl->ctx : nullptr)
    REQUIRE(gl && a && axis >= 0 && axis <= 2, "invalid arguments");
    checkBox(box);
    if (reset) resetCorr(gl, a->nlocal);
    REQUIRE(gl->corrSize >= a->nlocal + a->nghost, "ghost layer not initialised for these atoms");
    createAxis(gl, a, box, axis);
    MRMD_API_END
}

int mrmd_halo_update(mrmd_ghost_layer* gl, mrmd_atoms* a, const mrmd_box* box)
{
    MRMD_API_BEGIN(gl ? gl->ctx : nullptr)
    REQUIRE(gl && a, "invalid arguments");
    checkBox(box);
    REQUIRE(gl->corrSize >= a->nlocal + a->nghost, "correspondingRealAtom shorter than the atoms");
    if (a->nghost == 0) return MRMD_OK;
    KTimer t(gl->ctx, "halo_update");
    hipLaunchKernelGGL(k_ghostUpdate<false>, dim3(gridFor(a->nghost)), dim3(BLOCK), 0, gl->ctx->stream, a->pos.p,
                       gl->corr.p, a->nlocal, a->nlocal + a->nghost, makeSub(box), nullptr, nullptr);
    checkLaunch();
    MRMD_API_END
}

int mrmd_halo_fold_forces(mrmd_ghost_layer* gl, mrmd_atoms* a)
{
    MRMD_API_BEGIN(gl ? gl->ctx : nullptr)
    REQUIRE(gl && a, "invalid arguments");
    REQUIRE(gl->corrSize >= a->nlocal + a->nghost, "correspondingRealAtom shorter than the atoms");
    if (a->nghost == 0) return MRMD_OK;
    KTimer t(gl->ctx, "halo_fold");
    hipLaunchKernelGGL(k_ghostFold, dim3(gridFor(a->nghost)), dim3(BLOCK), 0, gl->ctx->stream, a->force.p, a->cap,
                       gl->corr.p, a->nlocal, a->nlocal + a->nghost);
    checkLaunch();
    MRMD_API_END
}

int mrmd_halo_get_corresponding(const mrmd_ghost_layer* gl, int64_t* host, int64_t n)
{
    MRMD_API_BEGIN(gl ? gl->ctx : nullptr)
    REQUIRE(gl && host && n >= 0 && n <= gl->corrSize, "invalid arguments (n=%lld, size=%lld)", (long long)n,
            gl ? (long long)gl->corrSize : -1LL);
    std::vector<int32_t> tmp(static_cast<size_t>(n));
    if (n)
    {
        HIPCHECK(hipMemcpyAsync(tmp.data(), gl->corr.p, size_t(n) * 4, hipMemcpyDeviceToHost, gl->ctx->stream));
        sync(gl->ctx);
    }
    for (int64_t i = 0; i < n; ++i) host[i] = tmp[size_t(i)];
    MRMD_API_END
}

int mrmd_halo_set_corresponding(mrmd_ghost_layer* gl, const int64_t* host, int64_t n)
{
    MRMD_API_BEGIN(gl ? gl->ctx : nullptr)
    REQUIRE(gl && host && n >= 0, "invalid arguments");
    std::vector<int32_t> tmp(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) tmp[size_t(i)] = int32_t(host[i]);
    gl->corr.reserve(gl->ctx, std::max<int64_t>(n, 1), 0);
    if (n)
    {
        HIPCHECK(hipMemcpyAsync(gl->corr.p, tmp.data(), size_t(n) * 4, hipMemcpyHostToDevice, gl->ctx->stream));
        sync(gl->ctx);
    }
    gl->corrSize = n;
    MRMD_API_END
}

int mrmd_multires_wrap(mrmd_molecules* m, mrmd_atoms* a, const mrmd_box* box)
{
    MRMD_API_BEGIN(m ? m->ctx : nullptr)
    REQUIRE(m && a, "invalid arguments");
    checkBox(box);
    if (m->nlocal == 0) return MRMD_OK;
    hipLaunchKernelGGL(k_molWrap, dim3(gridFor(m->nlocal)), dim3(BLOCK), 0, m->ctx->stream, m->view(), a->view(),
                       m->nlocal, makeSub(box));
    checkLaunch();
    MRMD_API_END
}

int mrmd_multires_halo_create(mrmd_ghost_layer* gl, mrmd_molecules* m, mrmd_atoms* a, const mrmd_box* box)
{
    MRMD_API_BEGIN(gl ? gl->ctx : nullptr)
    REQUIRE(gl && m && a, "invalid arguments");
    checkBox(box);
    m->nghost = 0;
    ++m->layoutVersion;
    resetCorr(gl, a->nlocal);
    a->nghost = 0;
    molCreateAll(gl, m, a, box);
    MRMD_API_END
}

int mrmd_multires_halo_create_axis(mrmd_ghost_layer* gl, mrmd_molecules* m, mrmd_atoms* a, const mrmd_box* box,
                                   int axis, int reset)
{
    MRMD_API_BEGIN(gl ? gl->ctx : nullptr)
    REQUIRE(gl && m && a && axis >= 0 && axis <= 2, "invalid arguments");
    checkBox(box);
    if (reset) resetCorr(gl, a->nlocal);
    molCreateAxis(gl, m, a, box, axis);
    MRMD_API_END
}

}  // extern "C"
