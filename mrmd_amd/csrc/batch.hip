// Ensemble engine: many independent periodic replicas (plain LJ under the Langevin
// thermostat, or monatomic LJ / ideal-gas AdResS slabs) advanced together on one device.
//
// Layout.  One arena holds every replica's atoms (pos double4, vel and force
// component-major); replica r owns atoms [a0, a0 + n), sorted slots [a0, a0 + n), its own
// periodic cell grid (cells of edge >= r_n, at least two per axis) and its own tiles.  Each
// phase of a step is one launch over all replicas, and no value of one replica depends on
// another: the step's decisions (rebuild or not, sampling, thermodynamic-force updates) are
// taken on the device per replica, so a block of steps is captured once as a hipGraph.
//
// Periodicity without ghost records.  A tile (a run of K cells along z in up to 2 x 2
// neighbouring (i, j) columns) stages the (2 + 2) x (2 + 2) columns around it x (K + 2)
// cells in LDS; each staged piece carries the periodic image shift that brings it next to
// the tile, applied as x + L or x - L, i.e. exactly the position a ghost record would have,
// so pair distances equal those of the ghost-layer path bit for bit.  Two cells per axis
// are enough: the images of a column then come from different stencil offsets.  Records
// are z-sorted inside each cell, so a staged column run is z-sorted and the list build cuts
// each row's candidates in a column to a z window by binary search, as the single-system
// build does.
//
// Determinism.  Per-replica sums are ordered (per-tile partials, then one pass per replica);
// the compensation-energy samples of LJ_IdealGas are accumulated in 64-bit fixed point
// (scale 2^32), whose integer atomics give the same total in any order.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "detmath.hpp"
#include "nlist.hpp"
#include "physics.hpp"

using namespace mrmd_hip;

namespace
{
constexpr int B_COLS = 2;          // tile columns per axis
constexpr int B_STEN = B_COLS + 2; // staged columns per axis
constexpr int B_RANGES = B_STEN * B_STEN * 3;  // staged columns, up to three image pieces each
constexpr int B_SELF = B_COLS * B_COLS;
constexpr int B_SPAN = 18;         // cells of a tile's column run (K + 2), so K <= 16
constexpr int B_PAD = 4;           // staged records past the neighbourhood a candidate group may read
constexpr int COMP_BINS = 200;     // LJ_IdealGas compensation bins (LJ_IdealGas.hpp:54)
constexpr double COMP_SCALE = 4294967296.0;
constexpr int64_t SAMPLING_INTERVAL = 200;    // LJ_IdealGas defaults (LJ_IdealGas.hpp:56-57)
constexpr int64_t COMP_UPDATE_INTERVAL = 20000;
constexpr int MAX_BINS = 4096;     // thermodynamic-force bins of one replica

struct RepD
{
    int64_t a0, n;
    int32_t c0, nc[3];
    int32_t t0, ntiles, K, tilesPerCol;
    double L[3], dxc[3], rdxc[3];
    double dt, zeta, temperature, mass, halfSkin;
    double damping, sigma;  ///< O-step factors of the replica's mass, computed on the host
    uint64_t seed;
    int32_t adress, tfEvery;
    mrmd_slab slab;
    int32_t h0, nbins;
    double binWidth, invBinWidth, binVolume, forceFactor, tfSigma;
    int32_t tfDelta, tfSym, tfPeriodic, cmp0;
};

// static geometry of a tile: columns i0 .. i0 + ni - 1 by j0 .. j0 + nj - 1, cells k0 .. k1 - 1
struct TileG
{
    int32_t rep, i0, j0, ni, nj, k0, k1;
};

// slot ranges of a tile, rebuilt with the replica's binning: the staged pieces (column by
// column, (ni + 2) x (nj + 2) columns, i-major), and the rows, self column after self column
struct TileB
{
    int32_t nr, staged, nself, nrows;
    int32_t begin[B_RANGES], base[B_RANGES];
    int8_t sx[B_RANGES], sy[B_RANGES], sz[B_RANGES];
    int32_t colBase[B_STEN * B_STEN];
    int32_t selfPiece[B_SELF], selfBegin[B_SELF], selfPrefix[B_SELF + 1];
};

// sorted slot of the t-th row of a tile, and its self column
__device__ inline int bTileRow(const TileB& ti, int t, int& sc)
{
    int r = 0;
    while (r + 1 < ti.nself && ti.selfPrefix[r + 1] <= t) ++r;
    sc = r;
    return ti.selfBegin[r] + (t - ti.selfPrefix[r]);
}

struct BS
{
    const RepD* reps;
    int32_t nrep;
    int64_t natoms, cap, ncells, ntiles;
    double4* pos;
    double* vel;
    double* force;
    const int32_t* atomRep;
    const int32_t* cellRep;
    int32_t* cellOf;
    int32_t* counts;
    int32_t* offsets;
    int32_t* fill;
    int32_t* sorted;
    int32_t* id;       ///< the atom's index in the ensemble as given (atoms move to slot order)
    int32_t* tmpA;     ///< permutation scratch (given indices, cells)
    int32_t* tmpB;
    double* tmpVel;
    double4* sortedPos;
    double* slotLam;
    double* slotGrad;
    const TileG* tg;
    TileB* tiles;
    uint16_t* nbr;
    int32_t* cnt;
    int64_t strideS;
    int32_t capList, stageCap;  ///< list width; staged records that fit the launch's LDS
    double* dispStep;
    double* dispAcc;
    int32_t* flag;
    int64_t* rebuilds;
    double* ePart;
    double* tfDensity;
    double* tfForce;
    int64_t* tfSamples;
    unsigned long long* comp;
    unsigned long long* compCount;
    double* meanComp;
    int64_t* step;
    int32_t* err;
    Pre lj;
    double rcSqr, rnSqr;
    int32_t tileBase;  ///< first tile of the launch (the force launches take plain and AdResS tiles apart)
};

__device__ inline int cellIndex(const RepD& R, int ix, int iy, int iz) { return R.c0 + (ix * R.nc[1] + iy) * R.nc[2] + iz; }

__device__ inline int locate(double x, double rdx, int n)
{
    int c = int(floor(x * rdx));
    c = c < 0 ? 0 : c;
    return c >= n ? n - 1 : c;
}

// ---------------------------------------------------------------------------------------
// integration

// VelocityVerletLangevinThermostat::preForceIntegrate (VelocityVerletLangevinThermostat.hpp:63-136);
// POST: the previous step's post-force integration (k_bPost) first, in the same pass: the
// thermodynamic force added to F_x, the kick, then this step's kick with the same force,
// the expressions and order of the two kernels
template <bool POST>
__global__ __launch_bounds__(BLOCK) void k_bPre(BS b)
{
    __shared__ WaveDraws draws[BLOCK / WAVE];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = i < b.natoms;
    const int r = valid ? b.atomRep[i] : -1;
    double d2 = 0.0;
    const RepD* R = valid ? &b.reps[r] : nullptr;
    const int64_t cap = b.cap;
    double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
    double vx = 0.0, vy = 0.0, vz = 0.0, ox = 0.0, oy = 0.0, oz = 0.0, dtHalf = 0.0;
    if (valid)
    {
        const double dt = R->dt, m = R->mass;
        dtHalf = 0.5 * dt;
        p = b.pos[i];
        vx = b.vel[i];
        vy = b.vel[cap + i];
        vz = b.vel[2 * cap + i];
        const double dtfm = dtHalf / m;
        double fx = b.force[i];
        const double fy = b.force[cap + i], fz = b.force[2 * cap + i];
        if (POST)
        {
            if (R->adress)
            {
                const int64_t bin = getBin(p.x, 0.0, R->invBinWidth, R->nbins);
                if (bin != -1) fx += b.tfForce[R->h0 + bin];
            }
            const double dtfmPost = (0.5 * R->dt) / R->mass;
            vx += dtfmPost * fx;
            vy += dtfmPost * fy;
            vz += dtfmPost * fz;
        }
        vx += dtfm * fx;
        vy += dtfm * fy;
        vz += dtfm * fz;
        ox = p.x;
        oy = p.y;
        oz = p.z;
        p.x += dtHalf * vx;
        p.y += dtHalf * vy;
        p.z += dtHalf * vz;
    }
    // the noise, one replica of the wave at a time (nearly always one): its seed is then
    // wave-uniform, so the Philox key stays in scalar registers
    double xi[3] = {0.0, 0.0, 0.0};
    const uint64_t id = valid ? uint64_t(b.id[i] - R->a0) : 0;
    for (unsigned long long pending = __ballot(valid); pending;)
    {
        const int rr = __builtin_amdgcn_readfirstlane(__shfl(r, __builtin_ctzll(pending), WAVE));
        const bool mine = valid && r == rr;
        double xt[3];
        waveNormals3(b.reps[rr].seed, mine ? id : 0, uint64_t(*b.step), mine, draws[threadIdx.x / WAVE], xt);
        if (mine)
            for (int d = 0; d < 3; ++d) xi[d] = xt[d];
        pending &= ~__ballot(mine);
    }
    if (valid)
    {
        const double damping = R->damping, sigma = R->sigma;
        vx *= damping;
        vy *= damping;
        vz *= damping;
        vx += sigma * xi[0];
        vy += sigma * xi[1];
        vz += sigma * xi[2];
        p.x += dtHalf * vx;
        p.y += dtHalf * vy;
        p.z += dtHalf * vz;
        b.pos[i] = p;
        b.vel[i] = vx;
        b.vel[cap + i] = vy;
        b.vel[2 * cap + i] = vz;
        const double dx = ox - p.x, dy = oy - p.y, dz = oz - p.z;
        d2 = dx * dx + dy * dy + dz * dz;
    }
    // the maximum aggregated over the workgroup's waves of one replica each (one atomic per
    // run of equal replicas among them); a wave of several replicas posts per lane
    __shared__ double wmax[BLOCK / WAVE];
    __shared__ int wrep[BLOCK / WAVE];
    const int r0 = __shfl(r, 0, WAVE);
    const bool single = __all(r == r0 || r < 0) && r0 >= 0;
    const double mx = waveMax(d2);
    if (threadIdx.x % WAVE == 0)
    {
        wmax[threadIdx.x / WAVE] = mx;
        wrep[threadIdx.x / WAVE] = single ? r0 : -1;
    }
    if (!single && valid) atomicMaxNonNeg(&b.dispStep[r], d2);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        int cur = -1;
        double m = 0.0;
        for (int w = 0; w < BLOCK / WAVE; ++w)
        {
            if (wrep[w] < 0) continue;
            if (wrep[w] != cur)
            {
                if (cur >= 0) atomicMaxNonNeg(&b.dispStep[cur], m);
                cur = wrep[w];
                m = wmax[w];
            }
            else
                m = fmax(m, wmax[w]);
        }
        if (cur >= 0) atomicMaxNonNeg(&b.dispStep[cur], m);
    }
}

// rebuild criterion per replica (examples/02_LennardJones_NVE/02_LennardJones_NVE.cpp:138-165)
__global__ void k_bCheck(BS b)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= b.nrep) return;
    const double d = sqrt(b.dispStep[r]);
    b.dispStep[r] = 0.0;
    const double acc = b.dispAcc[r] + d;
    if (acc >= b.reps[r].halfSkin)
    {
        b.flag[r] = 1;
        b.dispAcc[r] = 0.0;
        b.rebuilds[r] += 1;
    }
    else
    {
        b.flag[r] = 0;
        b.dispAcc[r] = acc;
    }
}

// PeriodicMapping::mapIntoDomain (PeriodicMapping.cpp:30-58) on the box [0, L)
__global__ void k_bWrap(BS b)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= b.natoms) return;
    const int r = b.atomRep[i];
    if (!b.flag[r]) return;
    const RepD& R = b.reps[r];
    double4 p = b.pos[i];
    double x[3] = {p.x, p.y, p.z};
    for (int d = 0; d < 3; ++d)
    {
        if (R.L[d] <= x[d])
        {
            x[d] -= R.L[d];
            x[d] = fmax(x[d], 0.0);
        }
        if (x[d] < 0.0)
        {
            x[d] += R.L[d];
            if (R.L[d] <= x[d]) x[d] = 0.0;
        }
    }
    p.x = x[0];
    p.y = x[1];
    p.z = x[2];
    b.pos[i] = p;
}

// ---------------------------------------------------------------------------------------
// binning of flagged replicas (stable counting sort per replica)

__global__ void k_bZeroCells(BS b)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= b.ncells || !b.flag[b.cellRep[c]]) return;
    b.counts[c] = 0;
    b.fill[c] = 0;
}

// atoms sit in cell order after their replica's last rebuild, so a wave's atoms share cells
// over long runs: one atomic per run (laneRun; valid lanes are a prefix of the wave)
__global__ void k_bCount(BS b)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int c = -1;
    if (i < b.natoms)
    {
        const int r = b.atomRep[i];
        if (b.flag[r])
        {
            const RepD& R = b.reps[r];
            const double4 p = b.pos[i];
            c = cellIndex(R, locate(p.x, R.rdxc[0], R.nc[0]), locate(p.y, R.rdxc[1], R.nc[1]),
                          locate(p.z, R.rdxc[2], R.nc[2]));
            b.cellOf[i] = c;
        }
    }
    const Run run = laneRun(c, i < b.natoms);
    if (c >= 0 && run.head == int(threadIdx.x) % WAVE) atomicAdd(&b.counts[c], run.len);
}

// one workgroup (BLOCK threads) per replica: offsets = a0 + exclusive scan of its counts
__global__ __launch_bounds__(BLOCK) void k_bScan(BS b)
{
    __shared__ int32_t lds[BLOCK / WAVE];
    const int r = blockIdx.x;
    if (!b.flag[r]) return;
    const RepD& R = b.reps[r];
    const int ncell = R.nc[0] * R.nc[1] * R.nc[2];
    int32_t carry = int32_t(R.a0);
    for (int base = 0; base < ncell; base += BLOCK)
    {
        const int c = base + int(threadIdx.x);
        const int32_t v = c < ncell ? b.counts[R.c0 + c] : 0;
        int32_t tot;
        const int32_t ex = blockExclusiveScan(v, lds, &tot);
        if (c < ncell) b.offsets[R.c0 + c] = carry + ex;
        carry += tot;
    }
}

__global__ void k_bFill(BS b)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = i < b.natoms;
    const int c = valid && b.flag[b.atomRep[i]] ? b.cellOf[i] : -1;
    const Run run = laneRun(c, valid);
    const int lane = int(threadIdx.x) % WAVE;
    int start = 0;
    if (c >= 0 && run.head == lane) start = atomicAdd(&b.fill[c], run.len);
    start = __shfl(start, run.head, WAVE);
    if (c >= 0) b.sorted[b.offsets[c] + start + (lane - run.head)] = int32_t(i);
}

// ascending z inside each cell of a flagged replica, the atom's given index breaking ties
// (the fill order of the atomics is arbitrary, and atoms move): one group of 32 lanes per
// cell (sortCellByZ)
__global__ __launch_bounds__(BLOCK) void k_bSortCells(BS b)
{
    const int64_t c = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / SORT_GROUP;
    if (c >= b.ncells || !b.flag[b.cellRep[c]]) return;
    sortCellByZ<true>(b.sorted + b.offsets[c], b.counts[c], b.pos, int(threadIdx.x % SORT_GROUP), b.id);
}

// A flagged replica's atoms move to its new slot order (position, velocity, given index and
// cell; forces are written by the step's force pass before they are read), so that the
// per-step slot gather, the force stores and the next binning run in atom order.  Pass 1
// gathers into scratch (the slot positions serve for the positions), pass 2 copies back and
// makes the slot map the identity.  Results do not depend on where an atom sits: its noise
// stream follows its given index, and cell-order ties break by it.
__global__ void k_bPermute1(BS b)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= b.natoms || !b.flag[b.atomRep[s]]) return;
    const int32_t i = b.sorted[s];
    b.sortedPos[s] = b.pos[i];
    for (int d = 0; d < 3; ++d) b.tmpVel[d * b.cap + s] = b.vel[d * b.cap + i];
    b.tmpA[s] = b.id[i];
    b.tmpB[s] = b.cellOf[i];
}

__global__ void k_bPermute2(BS b)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= b.natoms || !b.flag[b.atomRep[s]]) return;
    b.pos[s] = b.sortedPos[s];
    for (int d = 0; d < 3; ++d) b.vel[d * b.cap + s] = b.tmpVel[d * b.cap + s];
    b.id[s] = b.tmpA[s];
    b.cellOf[s] = b.tmpB[s];
    b.sorted[s] = int32_t(s);
}

// slot ranges, image shifts and LDS bases of every tile of a flagged replica: the staged
// columns i-major, each the pieces of its z-run [k0 - 1, k1 + 1) (below 0 the image -1,
// from nz the image +1), and the rows of the self columns (their unshifted pieces)
// one wave per tile (four per workgroup): lane L takes staged column L of the
// (ni + 2) x (nj + 2) columns (i-major, the serial order of the layout), its up to three
// pieces and, for a self column, its rows; wave prefix sums place the pieces, the LDS bases
// and the rows in that order
__global__ __launch_bounds__(BLOCK) void k_bTiles(BS b)
{
    const int64_t t = int64_t(blockIdx.x) * (BLOCK / WAVE) + threadIdx.x / WAVE;
    if (t >= b.ntiles) return;
    const TileG G = b.tg[t];
    if (!b.flag[G.rep]) return;
    const RepD& R = b.reps[G.rep];
    const int nx = R.nc[0], ny = R.nc[1], nz = R.nc[2];
    TileB& ti = b.tiles[t];
    const int lane = int(threadIdx.x) % WAVE;
    const int spanJ = G.nj + 2, ncol = (G.ni + 2) * spanJ;
    const bool col = lane < ncol;
    const int sa = lane / spanJ, sb = lane % spanJ;
    int ii = G.i0 - 1 + sa, jj = G.j0 - 1 + sb, sx = 0, sy = 0;
    if (ii < 0) ii += nx, sx = -1;
    else if (ii >= nx) ii -= nx, sx = 1;
    if (jj < 0) jj += ny, sy = -1;
    else if (jj >= ny) jj -= ny, sy = 1;
    const bool self = col && sa >= 1 && sa <= G.ni && sb >= 1 && sb <= G.nj;
    const int zs = G.k0 - 1, ze = G.k1 + 1;
    const int lo[3] = {zs < 0 ? nz - 1 : 0, zs < 0 ? 0 : zs, 0};
    const int hi[3] = {zs < 0 ? nz : 0, ze < nz ? ze : nz, ze > nz ? ze - nz : 0};
    int np = 0, len[3] = {0, 0, 0}, beg[3] = {0, 0, 0}, mainPiece = 0, rows = 0, rowBegin = 0;
    if (col)
        for (int pc = 0; pc < 3; ++pc)
        {
            if (hi[pc] <= lo[pc]) continue;
            const int cf = cellIndex(R, ii, jj, lo[pc]), cl = cellIndex(R, ii, jj, hi[pc] - 1);
            beg[pc] = b.offsets[cf];
            len[pc] = b.offsets[cl] + b.counts[cl] - beg[pc];
            if (pc == 1) mainPiece = np;
            ++np;
        }
    if (self)
    {
        const int rf = cellIndex(R, ii, jj, G.k0), rl = cellIndex(R, ii, jj, G.k1 - 1);
        rowBegin = b.offsets[rf];
        rows = b.offsets[rl] + b.counts[rl] - rowBegin;
    }
    const int staged = len[0] + len[1] + len[2];
    const int pieceBase = waveInclusiveScan(np) - np;
    const int stagedBase = waveInclusiveScan(staged) - staged;
    const int selfIdx = waveInclusiveScan(self ? 1 : 0) - (self ? 1 : 0);
    const int rowsBase = waveInclusiveScan(rows) - rows;
    if (col)
    {
        ti.colBase[lane] = stagedBase;
        const int sh[3] = {-1, 0, 1};
        int nr = pieceBase, base = stagedBase;
        for (int pc = 0; pc < 3; ++pc)
        {
            if (hi[pc] <= lo[pc]) continue;
            ti.begin[nr] = beg[pc];
            ti.base[nr] = base;
            ti.sx[nr] = int8_t(sx);
            ti.sy[nr] = int8_t(sy);
            ti.sz[nr] = int8_t(sh[pc]);
            base += len[pc];
            ++nr;
        }
    }
    if (self)
    {
        ti.selfPiece[selfIdx] = pieceBase + mainPiece;
        ti.selfBegin[selfIdx] = rowBegin;
        ti.selfPrefix[selfIdx] = rowsBase;
    }
    const int nrTot = __shfl(pieceBase + np, WAVE - 1, WAVE), stTot = __shfl(stagedBase + staged, WAVE - 1, WAVE);
    const int nself = __shfl(selfIdx + (self ? 1 : 0), WAVE - 1, WAVE), nrows = __shfl(rowsBase + rows, WAVE - 1, WAVE);
    if (lane == 0)
    {
        ti.selfPrefix[nself] = nrows;
        ti.nr = nrTot;
        ti.staged = stTot;
        ti.nself = nself;
        ti.nrows = nrows;
    }
}

// ---------------------------------------------------------------------------------------
// per-step slot data: current positions in sorted order with the slab weights, and the
// density sample of the thermodynamic force (AxialDensityProfile.cpp:21-51)
__global__ void k_bGather(BS b)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = s < b.natoms;
    int64_t key = -1;  // density bin of the atom's replica, -1 for none
    if (valid)
    {
        const int32_t i = b.sorted[s];
        const double4 p = b.pos[i];
        const RepD& R = b.reps[b.atomRep[i]];
        double lam = 1.0, mod = 1.0, g = 0.0;
        if (R.adress)
        {
            slabEval(R.slab, p.x, lam, mod, g);
            const int64_t bin = getBin(p.x, 0.0, R.invBinWidth, R.nbins);
            if (bin != -1) key = R.h0 + bin;
        }
        b.sortedPos[s] = make_double4(p.x, p.y, p.z, mod);
        // lambda and its gradient are read for hybrid records only (0 < lambda_mod < 1)
        if (mod > 0.0 && mod < 1.0)
        {
            b.slotLam[s] = lam;
            b.slotGrad[s] = g;
        }
    }
    // a wave's atoms (a few cells of one column, so one x-cell) fall into a few x-bins: one
    // atomic per distinct bin of the wave (counts are integers, so the sums are exact in any
    // order)
    const int lane = int(threadIdx.x) % WAVE;
    const int k32 = int(key);
    for (unsigned long long pending = __ballot(k32 >= 0); pending;)
    {
        const int src = __builtin_ctzll(pending);
        const int k0 = __shfl(k32, src, WAVE);
        const unsigned long long m = __ballot(k32 == k0);
        if (lane == src) atomicAddD(&b.tfDensity[k0], double(__popcll(m)));
        pending &= ~m;
    }
}

// copy a tile's neighbourhood into LDS, shifting each piece by its periodic image: {x, y, z}
// records (XYZ) or whole double4 records (the AdResS forces read the slab weight in w)
// Returns whether a record this lane staged has a weight w > 0 (whole records only).
template <bool XYZ>
__device__ inline bool stageShifted(const TileB& ti, const double4* __restrict__ sp, double* st, const RepD& R)
{
    bool weighted = false;
    int r = 0;
    for (int q = threadIdx.x; q < ti.staged; q += blockDim.x)
    {
        while (r + 1 < ti.nr && ti.base[r + 1] <= q) ++r;
        double4 v = sp[ti.begin[r] + (q - ti.base[r])];
        v.x += double(ti.sx[r]) * R.L[0];
        v.y += double(ti.sy[r]) * R.L[1];
        v.z += double(ti.sz[r]) * R.L[2];
        if (XYZ)
        {
            st[3 * q] = v.x;
            st[3 * q + 1] = v.y;
            st[3 * q + 2] = v.z;
        }
        else
        {
            reinterpret_cast<double4*>(st)[q] = v;
            weighted = weighted || v.w > 0.0;
        }
    }
    return weighted;
}

// A tile's records by staged slot: from LDS when its neighbourhood fits the stage area, else
// straight from the slot-ordered records with the same periodic shift (the same arithmetic,
// so the same bits) -- a slower path that keeps tiles of crowded cells (lattice starts,
// density fluctuations) exact instead of dropping them.
struct LdsFetch3
{
    const double* st;
    __device__ double4 operator()(int q) const { return make_double4(st[3 * q], st[3 * q + 1], st[3 * q + 2], 1.0); }
};
struct LdsFetch4
{
    const double4* st;
    __device__ double4 operator()(int q) const { return st[q]; }
};
struct GlobalFetch
{
    const TileB* ti;
    const double4* sp;
    double L0, L1, L2;
    __device__ double4 operator()(int q) const
    {
        q = min(q, ti->staged - 1);  // a candidate group may read up to three records past a run
        int r = 0;
        while (r + 1 < ti->nr && ti->base[r + 1] <= q) ++r;
        double4 v = sp[ti->begin[r] + (q - ti->base[r])];
        v.x += double(ti->sx[r]) * L0;
        v.y += double(ti->sy[r]) * L1;
        v.z += double(ti->sz[r]) * L2;
        return v;
    }
};

// Full rows of one tile of a flagged replica, one lane per row, as the single-system build
// forms them: per stencil column the cells that Cabana's min-distance prune keeps (one run
// in z), cut by binary search to the z window of the row (records are z-sorted in a staged
// column), 32 candidates per pass in groups of four with the hit bits shifted in by a
// funnel shift, then the hits appended in slot order.  Entries are staged slots; rows are
// padded to whole chunks with the row's own slot (distance 0, which the forces skip).
template <class Fetch>
__device__ void bBuildRows(const BS& b, const TileB& ti, const TileG& G, const RepD& R, const int* cstart,
                           const Fetch& fetch)
{
    const int spanJ = G.nj + 2;
    const int64_t chunkStride = b.strideS * CHUNK;
    const double rsq = b.rnSqr;
    int localMax = 0;
    for (int t = threadIdx.x; t < ti.nrows; t += blockDim.x)
    {
        int sc;
        const int s = bTileRow(ti, t, sc);
        const int pr = ti.selfPiece[sc];
        const int qSelf = ti.base[pr] + (s - ti.begin[pr]);
        const double4 xp = fetch(qSelf);
        const int kz = (b.cellOf[b.sorted[s]] - R.c0) % R.nc[2];
        const int sa = sc / G.nj, sb = sc % G.nj;
        const int i = G.i0 + sa, j = G.j0 + sb;
        // per-axis distances from the row to the stencil cells' boxes (unwrapped indices) and
        // the lateral terms of the z window, shrunk by 1e-9 of a cell
        double rx2[3], ry2[3], wx2[3], wy2[3], rz2[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            const double cx = (double(i - 1 + a) + 0.5) * R.dxc[0];
            const double cy = (double(j - 1 + a) + 0.5) * R.dxc[1];
            const double cz = (double(kz - 1 + a) + 0.5) * R.dxc[2];
            double r = fabs(xp.x - cx) - 0.5 * R.dxc[0];
            r = r > 0.0 ? r : 0.0;
            rx2[a] = r * r;
            r = fabs(xp.y - cy) - 0.5 * R.dxc[1];
            r = r > 0.0 ? r : 0.0;
            ry2[a] = r * r;
            r = fabs(xp.z - cz) - 0.5 * R.dxc[2];
            r = r > 0.0 ? r : 0.0;
            rz2[a] = r * r;
            double e = fabs(xp.x - cx) - 0.5 * R.dxc[0] - 1e-9 * R.dxc[0];
            e = e > 0.0 ? e : 0.0;
            wx2[a] = e * e;
            e = fabs(xp.y - cy) - 0.5 * R.dxc[1] - 1e-9 * R.dxc[1];
            e = e > 0.0 ? e : 0.0;
            wy2[a] = e * e;
        }
        uint16_t* const rowBase = b.nbr + int64_t(s) * CHUNK;
        int c = 0;
#pragma unroll
        for (int da = -1; da <= 1; ++da)
#pragma unroll
            for (int db = -1; db <= 1; ++db)
            {
                // z index of the staged column: cell k of the run is entry k - (k0 - 1)
                const int* cs = cstart + ((sa + 1 + da) * spanJ + (sb + 1 + db)) * (B_SPAN + 1) - (G.k0 - 1);
                const double lat = rx2[da + 1] + ry2[db + 1];
                int kLo = INT_MAX, kHi = INT_MIN;
#pragma unroll
                for (int dc = 0; dc < 3; ++dc)
                    if (lat + rz2[dc] <= rsq)
                    {
                        kLo = min(kLo, kz - 1 + dc);
                        kHi = kz - 1 + dc;
                    }
                if (kLo > kHi) continue;
                const double zsq = rsq - (wx2[da + 1] + wy2[db + 1]) + 1e-12 * rsq;
                int lo = cs[kLo], hi = cs[kLo + 1];
                while (lo < hi)
                {
                    const int m = (lo + hi) >> 1;
                    const double dz = xp.z - fetch(m).z;
                    if (dz > 0.0 && dz * dz > zsq) lo = m + 1;
                    else hi = m;
                }
                const int wlo = lo;
                lo = max(wlo, cs[kHi]);
                hi = cs[kHi + 1];
                while (lo < hi)
                {
                    const int m = (lo + hi) >> 1;
                    const double dz = xp.z - fetch(m).z;
                    if (dz < 0.0 && dz * dz > zsq) hi = m;
                    else lo = m + 1;
                }
                const int whi = lo;
                for (int qb = wlo; qb < whi; qb += 32)
                {
                    uint32_t miss = 0;
                    const int nq = min(32, whi - qb);
                    int m = 0;
                    for (; m < nq; m += 4)
                    {
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                        {
                            const double4 xn = fetch(qb + m + u);
                            const double dx = xp.x - xn.x, dy = xp.y - xn.y, dz = xp.z - xn.z;
                            const double r2 = dx * dx + dy * dy + dz * dz;
                            miss = __builtin_amdgcn_alignbit(miss, uint32_t(__double2hiint(rsq - r2)), 31u);
                        }
                    }
                    uint32_t hits = ~miss << (32 - m);
                    hits &= ~((1u << (32 - nq)) - 1u);
                    if (qSelf >= qb && qSelf < qb + nq) hits &= ~(0x80000000u >> (qSelf - qb));
                    while (hits)
                    {
                        const int u = __builtin_clz(hits);
                        hits ^= 0x80000000u >> u;
                        if (c < b.capList) rowBase[int64_t(c >> 3) * chunkStride + (c & 7)] = uint16_t(qb + u);
                        ++c;
                    }
                }
            }
        for (int k = c; k < ((c + CHUNK - 1) / CHUNK) * CHUNK && k < b.capList; ++k)
            rowBase[int64_t(k >> 3) * chunkStride + (k & 7)] = uint16_t(qSelf);
        b.cnt[s] = min(c, b.capList);  // an overflow is reported through err; reads stay in the list
        localMax = max(localMax, c);
    }
    if (localMax > b.capList) atomicOr(b.err, 1);
}

// full-row list of every tile of a flagged replica
__global__ __launch_bounds__(BLOCK) void k_bBuild(BS b)
{
    extern __shared__ double stb[];
    __shared__ TileB ti;
    __shared__ int cstart[B_STEN * B_STEN * (B_SPAN + 1)];
    // plain block order: when only some replicas rebuild, their tiles are contiguous, and the
    // hardware's round-robin of consecutive workgroups over the XCDs spreads them over all
    // eight (grouping consecutive tiles on one XCD would put a whole replica on one)
    const unsigned t = blockIdx.x;
    const TileG G = b.tg[t];
    if (!b.flag[G.rep]) return;
    if (threadIdx.x == 0) ti = b.tiles[t];
    __syncthreads();
    if (ti.nrows == 0) return;
    const RepD R = b.reps[G.rep];
    if (R.adress)
    {
        // an AdResS tile whose staged columns all lie out of reach of the weighted region
        // (the slab's atomistic and hybrid x band widened by the displacement that triggers
        // a rebuild; weights come from the records' own x in [0, L)) keeps no pair before its
        // next rebuild: empty rows, and the force kernel finds every weight 0
        const double reach = 0.5 * R.slab.atomistic_diameter + R.slab.hybrid_diameter + R.halfSkin;
        const double lo = R.slab.center[0] - reach, hi = R.slab.center[0] + reach;
        bool near = false;
        for (int sa = 0; sa < G.ni + 2; ++sa)
        {
            int ii = G.i0 - 1 + sa;
            ii = ii < 0 ? ii + R.nc[0] : (ii >= R.nc[0] ? ii - R.nc[0] : ii);
            const double xlo = ii == 0 ? -HUGE_VAL : double(ii) * R.dxc[0] - 1e-9 * R.dxc[0];
            const double xhi = ii == R.nc[0] - 1 ? HUGE_VAL : double(ii + 1) * R.dxc[0] + 1e-9 * R.dxc[0];
            near = near || !(xhi < lo || xlo > hi);
        }
        if (!near)
        {
            for (int t = threadIdx.x; t < ti.nrows; t += blockDim.x)
            {
                int sc;
                b.cnt[bTileRow(ti, t, sc)] = 0;
            }
            return;
        }
    }
    const bool fits = ti.staged <= b.stageCap;
    if (fits) stageShifted<true>(ti, b.sortedPos, stb, R);
    // LDS slot of the first record of cell k0 - 1 + z of each staged column
    const int nzr = G.k1 - G.k0 + 2;
    const int spanJ = G.nj + 2, ncol = (G.ni + 2) * spanJ;
    for (int col = threadIdx.x; col < ncol; col += blockDim.x)
    {
        int ii = G.i0 - 1 + col / spanJ, jj = G.j0 - 1 + col % spanJ;
        ii = ii < 0 ? ii + R.nc[0] : (ii >= R.nc[0] ? ii - R.nc[0] : ii);
        jj = jj < 0 ? jj + R.nc[1] : (jj >= R.nc[1] ? jj - R.nc[1] : jj);
        int run = ti.colBase[col];
        for (int z = 0; z < nzr; ++z)
        {
            int kz = G.k0 - 1 + z;
            kz = kz < 0 ? kz + R.nc[2] : (kz >= R.nc[2] ? kz - R.nc[2] : kz);
            cstart[col * (B_SPAN + 1) + z] = run;
            run += b.counts[cellIndex(R, ii, jj, kz)];
        }
        cstart[col * (B_SPAN + 1) + nzr] = run;
    }
    __syncthreads();
    if (fits)
        bBuildRows(b, ti, G, R, cstart, LdsFetch3{stb});
    else
        bBuildRows(b, ti, G, R, cstart, GlobalFetch{&ti, b.sortedPos, R.L[0], R.L[1], R.L[2]});
}

// ---------------------------------------------------------------------------------------
// forces: weighted LJ (LJ_IdealGas.cpp:80-200 with one atom per molecule; weight 1 for
// plain replicas, which reduces to LennardJones.hpp:168-197), drift and compensation terms

// W: weighted (AdResS replicas: slab weights in w, the unweighted energy e0 for the drift);
// plain replicas have weight 1 and need neither
template <bool CAP, bool W, class Fetch>
__device__ inline void chunkW(const BS& b, const double4& pi, const Fetch& fetch, const uint4 cur, double& fx,
                              double& fy, double& fz, double& ew, double& e0)
{
    const uint32_t wd[4] = {cur.x, cur.y, cur.z, cur.w};
    double dx[CHUNK], dy[CHUNK], dz[CHUNK], d2[CHUNK], lj[CHUNK], frac2[CHUNK];
#pragma unroll
    for (int u = 0; u < CHUNK; ++u)
    {
        const double4 pj = fetch(int((wd[u >> 1] >> ((u & 1) * 16)) & 0x7fffu));
        dx[u] = pi.x - pj.x;
        dy[u] = pi.y - pj.y;
        dz[u] = pi.z - pj.z;
        lj[u] = pj.w;
    }
#pragma unroll
    for (int u = 0; u < CHUNK; ++u) d2[u] = dx[u] * dx[u] + dy[u] * dy[u] + dz[u] * dz[u];
#pragma unroll
    for (int u = 0; u < CHUNK; ++u) frac2[u] = recipNewton(d2[u]);
    const Pre& p = b.lj;
    auto accumulate = [&](int u, double ff, double en) {
        if constexpr (W)
        {
            // both partners coarse-grained: no interaction (LJ_IdealGas.cpp:95)
            const bool in = d2[u] <= b.rcSqr && d2[u] > 0.0 && !(pi.w <= 0.0 && lj[u] <= 0.0);
            const double w = 0.5 * (pi.w + lj[u]);
            const double ffw = in ? ff * w : 0.0;
            fx += dx[u] * ffw;
            fy += dy[u] * ffw;
            fz += dz[u] * ffw;
            ew += in ? en * w : 0.0;
            e0 += in ? en : 0.0;
        }
        else
        {
            // padding entries are the row itself (distance 0)
            const bool in = d2[u] <= b.rcSqr && d2[u] > 0.0;
            const double ffw = in ? ff : 0.0;
            fx += dx[u] * ffw;
            fy += dy[u] * ffw;
            fz += dz[u] * ffw;
            ew += in ? en : 0.0;
        }
    };
    if constexpr (!CAP)
    {
#pragma unroll
        for (int u = 0; u < CHUNK; ++u)
        {
            const double frac6 = frac2[u] * frac2[u] * frac2[u];
            accumulate(u, frac6 * (p.ff1 * frac6 - p.ff2) * frac2[u], frac6 * (p.ef1 * frac6 - p.ef2) - p.shift);
        }
        return;
    }
    double ffv[CHUNK], env[CHUNK];
    bool capped = false;
#pragma unroll
    for (int u = 0; u < CHUNK; ++u)
    {
        const double frac6 = frac2[u] * frac2[u] * frac2[u];
        ffv[u] = frac6 * (p.ff1 * frac6 - p.ff2) * frac2[u];
        env[u] = frac6 * (p.ef1 * frac6 - p.ef2) - p.shift;
        capped = capped || (d2[u] < p.capDistSqr && d2[u] > 0.0);  // padding (the row itself) is not a pair
    }
    // the capped form (LennardJones.hpp:68-75) only for chunks where some lane of the wave
    // needs it
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
}

// the rows of one tile: weighted pair forces, drift and compensation terms; returns the
// lane's energy sum
template <bool CAP, bool W, class Fetch>
__device__ double bForceRows(const BS& b, const TileB& ti, const RepD& R, const Fetch& fetch)
{
    const int64_t step = *b.step;
    const bool sampling = W && R.adress && step % SAMPLING_INTERVAL == 0;
    double eTile = 0.0;
    for (int t = threadIdx.x; t < ti.nrows; t += blockDim.x)
    {
        int sc;
        const int s = bTileRow(ti, t, sc);
        const int c = b.cnt[s];
        const int32_t i = b.sorted[s];
        const double4 pi = b.sortedPos[s];
        double fx = 0.0, fy = 0.0, fz = 0.0, ew = 0.0, e0 = 0.0;
        const uint4* chunk = reinterpret_cast<const uint4*>(b.nbr) + s;
        const int nch = (c + CHUNK - 1) / CHUNK;
        uint4 cur = chunk[0];  // the first chunk column exists for every slot (read even when c = 0)
        for (int ch = 0; ch < nch; ++ch)
        {
            uint4 nxt = cur;
            if (ch + 1 < nch) nxt = chunk[int64_t(ch + 1) * b.strideS];
            chunkW<CAP, W>(b, pi, fetch, cur, fx, fy, fz, ew, e0);
            cur = nxt;
        }
        if (W && R.adress && pi.w > 0.0 && pi.w < 1.0)
        {
            // hybrid atom: drift -1/2 sum_j E_ij grad(lambda) and the mean compensation
            // energy (LJ_IdealGas.cpp:150-170, 208-224)
            const double g = b.slotGrad[s];
            fx += -(0.5 * e0) * g;
            const int64_t bin = getBin(b.slotLam[s], 0.0, double(COMP_BINS), COMP_BINS);
            if (bin != -1)
            {
                fx += b.meanComp[R.cmp0 + bin] * g;
                if (sampling)
                {
                    atomicAdd(&b.comp[R.cmp0 + bin], (unsigned long long)(long long)llrint(0.5 * e0 * COMP_SCALE));
                    atomicAdd(&b.compCount[R.cmp0 + bin], 1ull);
                }
            }
        }
        b.force[i] = fx;
        b.force[b.cap + i] = fy;
        b.force[2 * b.cap + i] = fz;
        eTile += ew;
    }
    return eTile;
}

// tiles [tileBase, tileBase + gridDim.x): the plain replicas' tiles (W = false, {x, y, z}
// staged) or the AdResS replicas' (W = true, whole records staged)
template <bool CAP, bool W>
__global__ __launch_bounds__(BLOCK) void k_bForce(BS b)
{
    extern __shared__ double stf[];
    __shared__ TileB ti;
    __shared__ double red[BLOCK / WAVE];
    const unsigned t = unsigned(b.tileBase) + xcdBlock();
    const TileG G = b.tg[t];
    if (threadIdx.x == 0) ti = b.tiles[t];
    __syncthreads();
    const bool fits = ti.staged <= b.stageCap;
    const RepD R = b.reps[G.rep];
    const bool weighted = fits ? stageShifted<!W>(ti, b.sortedPos, stf, R) : true;
    // an AdResS tile whose neighbourhood is all coarse-grained (every weight 0) has no pair
    // (LJ_IdealGas.cpp:95), hence zero forces, drift and energy: the rows get the +0.0 that
    // the pair loop would leave, without it
    if (W && !__syncthreads_or(weighted))
    {
        for (int t = threadIdx.x; t < ti.nrows; t += blockDim.x)
        {
            int sc;
            const int32_t i = b.sorted[bTileRow(ti, t, sc)];
            b.force[i] = 0.0;
            b.force[b.cap + i] = 0.0;
            b.force[2 * b.cap + i] = 0.0;
        }
        if (threadIdx.x == 0) b.ePart[t] = 0.0;
        return;
    }
    if (!W) __syncthreads();
    double eTile;
    if (!fits)
        eTile = bForceRows<CAP, W>(b, ti, R, GlobalFetch{&ti, b.sortedPos, R.L[0], R.L[1], R.L[2]});
    else if constexpr (W)
        eTile = bForceRows<CAP, W>(b, ti, R, LdsFetch4{reinterpret_cast<const double4*>(stf)});
    else
        eTile = bForceRows<CAP, W>(b, ti, R, LdsFetch3{stf});
    const double e = blockSum(0.5 * eTile, red);
    if (threadIdx.x == 0) b.ePart[t] = e;
}

// LJ_IdealGas::updateMeanCompensationEnergy (LJ_IdealGas.cpp:21-50) at the reference's
// update interval; one workgroup per replica
__global__ void k_bCompUpdate(BS b)
{
    const int r = blockIdx.x;
    const RepD& R = b.reps[r];
    if (!R.adress || *b.step % COMP_UPDATE_INTERVAL != 0) return;
    for (int k = threadIdx.x; k < COMP_BINS; k += blockDim.x)
    {
        const int64_t idx = R.cmp0 + k;
        const double cntv = double(b.compCount[idx]);
        if (cntv < 0.5) continue;
        const double sum = double((long long)b.comp[idx]) / COMP_SCALE;
        b.meanComp[idx] = (10.0 * b.meanComp[idx] + sum / cntv) / 11.0;
        b.comp[idx] = 0ull;
        b.compCount[idx] = 0ull;
    }
}

// ThermodynamicForce::update_if (ThermodynamicForce.hpp:124-152) every tfEvery steps with
// the hybrid region as predicate; one workgroup per replica
__global__ __launch_bounds__(BLOCK) void k_bTFUpdate(BS b)
{
    __shared__ double dens[MAX_BINS];
    __shared__ double smooth[MAX_BINS];
    const int r = blockIdx.x;
    const RepD& R = b.reps[r];
    if (!R.adress) return;
    __syncthreads();
    const int64_t samples = b.tfSamples[r] + 1;
    if (R.tfEvery <= 0 || (*b.step + 1) % R.tfEvery != 0)
    {
        if (threadIdx.x == 0) b.tfSamples[r] = samples;
        return;
    }
    const int nb = R.nbins;
    double* density = b.tfDensity + R.h0;
    double* force = b.tfForce + R.h0;
    for (int c = threadIdx.x; c < nb; c += blockDim.x) dens[c] = density[c];
    __syncthreads();
    if (R.tfSym)
    {
        for (int c = threadIdx.x; c < nb / 2; c += blockDim.x)
        {
            const double v = 0.5 * (dens[c] + dens[nb - 1 - c]);
            dens[c] = v;
            dens[nb - 1 - c] = v;
        }
        __syncthreads();
    }
    const double norm = 1.0 / (R.binVolume * double(samples));
    for (int c = threadIdx.x; c < nb; c += blockDim.x) dens[c] *= norm;
    __syncthreads();
    const HistD hd{0.0, R.binWidth, R.invBinWidth, nb, 1};
    for (int c = threadIdx.x; c < nb; c += blockDim.x)
        smooth[c] = smoothAt(dens, hd, c, 0, R.tfSigma, R.tfDelta, R.tfPeriodic);
    __syncthreads();
    mrmd_pred1 hy{};
    hy.kind = MRMD_PRED_SLAB_HY;
    hy.axis = 0;
    hy.a = R.slab.center[0];
    hy.b = 0.5 * R.slab.atomistic_diameter;
    hy.c = R.slab.hybrid_diameter;
    for (int c = threadIdx.x; c < nb; c += blockDim.x)
    {
        double gr = gradientAt(smooth, hd, c, 0, R.tfPeriodic) * R.forceFactor;
        const double x = (double(c) + 0.5) * R.binWidth;
        if (!evalPred1(hy, x, x, x)) gr = 0.0;
        force[c] -= gr;
        density[c] = 0.0;
    }
    if (threadIdx.x == 0) b.tfSamples[r] = 0;
}

// VelocityVerletLangevinThermostat::postForceIntegrate with the thermodynamic force
// (ThermodynamicForce::apply, ThermodynamicForce.hpp:98-122) added first
__global__ void k_bPost(BS b)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= b.natoms) return;
    const RepD& R = b.reps[b.atomRep[i]];
    double fx = b.force[i];
    if (R.adress)
    {
        const int64_t bin = getBin(b.pos[i].x, 0.0, R.invBinWidth, R.nbins);
        if (bin != -1)
        {
            fx += b.tfForce[R.h0 + bin];
            b.force[i] = fx;
        }
    }
    const double dtfm = (0.5 * R.dt) / R.mass;
    b.vel[i] += dtfm * fx;
    b.vel[b.cap + i] += dtfm * b.force[b.cap + i];
    b.vel[2 * b.cap + i] += dtfm * b.force[2 * b.cap + i];
}

__global__ void k_bStepEnd(BS b)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *b.step += 1;
}

// per-replica ordered sums: potential energy from the tile partials, kinetic energy
__global__ __launch_bounds__(BLOCK) void k_bEnergies(BS b, double* out)
{
    __shared__ double red[BLOCK / WAVE];
    const int r = blockIdx.x;
    const RepD& R = b.reps[r];
    double e = 0.0;
    for (int t = threadIdx.x; t < R.ntiles; t += blockDim.x) e += b.ePart[R.t0 + t];
    e = blockSum(e, red);
    double k = 0.0;
    for (int64_t i = R.a0 + threadIdx.x; i < R.a0 + R.n; i += blockDim.x)
    {
        const double vx = b.vel[i], vy = b.vel[b.cap + i], vz = b.vel[2 * b.cap + i];
        k += vx * vx + vy * vy + vz * vz;
    }
    k = blockSum(k, red);
    if (threadIdx.x == 0)
    {
        out[2 * r] = e;
        out[2 * r + 1] = 0.5 * R.mass * k;
    }
}

__global__ void k_bFillD(double* p, int64_t n, double v)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

__global__ void k_bIota(int32_t* p, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = int32_t(k);
}
}  // namespace

struct mrmd_batch
{
    CtxPtr ctx;
    std::vector<mrmd_replica> host;
    std::vector<RepD> reps;
    std::vector<TileG> tg;
    int64_t natoms = 0, cap = 0, ncells = 0, ntiles = 0, nbins = 0;
    double rc = 2.5, skin = 0.3;
    Pre lj{};
    bool cap_ = false;
    int targetStaged = 1400, block = 256;
    int stageCapPlain = 1600, stageCapAdress = 1600, stageCapBuild = 1600;  ///< staged records in LDS per launch
    int stageCapLimit = 0;  ///< 0: the stage area follows the expected neighbourhood size
    int64_t plainTiles = 0;  ///< tiles of the plain replicas come first, then the AdResS replicas' tiles
    int32_t capList = 0;
    int64_t steps = 0;

    DBuf<RepD> dreps;
    DBuf<double4> pos, sortedPos;
    DBuf<double> vel, force, slotLam, slotGrad;
    DBuf<int32_t> atomRep, cellRep, cellOf, counts, offsets, fill, sorted, cnt, flag, err;
    DBuf<int32_t> id, tmpA, tmpB;
    DBuf<double> tmpVel;
    DBuf<TileG> dtg;
    DBuf<TileB> tiles;
    DBuf<uint16_t> nbr;
    DBuf<double> dispStep, dispAcc, ePart, tfDensity, tfForce, meanComp, energies;
    DBuf<int64_t> rebuilds, tfSamples, step;
    DBuf<unsigned long long> comp, compCount;

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graphSteps = 0;

    ~mrmd_batch() { dropGraph(); }
    void dropGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graphSteps = 0;
    }
    BS state();
    void layoutTiles();
    void enqueueSteps(int64_t n);
};

BS mrmd_batch::state()
{
    BS b{};
    b.reps = dreps.p;
    b.nrep = int32_t(reps.size());
    b.natoms = natoms;
    b.cap = cap;
    b.ncells = ncells;
    b.ntiles = ntiles;
    b.pos = pos.p;
    b.vel = vel.p;
    b.force = force.p;
    b.atomRep = atomRep.p;
    b.cellRep = cellRep.p;
    b.cellOf = cellOf.p;
    b.counts = counts.p;
    b.offsets = offsets.p;
    b.fill = fill.p;
    b.sorted = sorted.p;
    b.id = id.p;
    b.tmpA = tmpA.p;
    b.tmpB = tmpB.p;
    b.tmpVel = tmpVel.p;
    b.sortedPos = sortedPos.p;
    b.slotLam = slotLam.p;
    b.slotGrad = slotGrad.p;
    b.tg = dtg.p;
    b.tiles = tiles.p;
    b.nbr = nbr.p;
    b.cnt = cnt.p;
    b.strideS = cap;
    b.capList = capList;
    b.stageCap = stageCapBuild;
    b.tileBase = 0;
    b.dispStep = dispStep.p;
    b.dispAcc = dispAcc.p;
    b.flag = flag.p;
    b.rebuilds = rebuilds.p;
    b.ePart = ePart.p;
    b.tfDensity = tfDensity.p;
    b.tfForce = tfForce.p;
    b.tfSamples = tfSamples.p;
    b.comp = comp.p;
    b.compCount = compCount.p;
    b.meanComp = meanComp.p;
    b.step = step.p;
    b.err = err.p;
    b.lj = lj;
    b.rcSqr = rc * rc;
    b.rnSqr = (rc + skin) * (rc + skin);
    return b;
}

// per-replica tile height from the target staged size, tile geometry, LDS capacity; the
// plain replicas' tiles first, then the AdResS replicas' (their force launches stage
// different records)
void mrmd_batch::layoutTiles()
{
    tg.clear();
    double maxStage[2] = {0.0, 0.0};
    int64_t t0 = 0;
    for (int adress = 0; adress < 2; ++adress)
    {
        if (adress == 1) plainTiles = t0;
        for (auto& R : reps)
        {
            if (R.adress != adress) continue;
            const double perCell = double(R.n) / double(R.nc[0] * R.nc[1] * R.nc[2]);
            const double cols = double(B_STEN * B_STEN);
            int K = int(std::lround(double(targetStaged) / (cols * std::max(perCell, 1e-3)))) - 2;
            K = std::max(1, std::min({K, R.nc[2], B_SPAN - 2}));
            R.K = K;
            R.tilesPerCol = (R.nc[2] + K - 1) / K;
            R.t0 = int32_t(t0);
            const int bi = (R.nc[0] + B_COLS - 1) / B_COLS, bj = (R.nc[1] + B_COLS - 1) / B_COLS;
            R.ntiles = bi * bj * R.tilesPerCol;
            for (int ib = 0; ib < bi; ++ib)
                for (int jb = 0; jb < bj; ++jb)
                    for (int kb = 0; kb < R.tilesPerCol; ++kb)
                        tg.push_back(TileG{int32_t(&R - reps.data()), ib * B_COLS, jb * B_COLS,
                                           std::min(B_COLS, R.nc[0] - ib * B_COLS),
                                           std::min(B_COLS, R.nc[1] - jb * B_COLS), kb * K,
                                           std::min((kb + 1) * K, R.nc[2])});
            t0 += R.ntiles;
            maxStage[adress] = std::max(maxStage[adress], cols * (K + 2) * perCell);
        }
    }
    ntiles = t0;
    // LDS capacity: the expected largest neighbourhood with room for density fluctuations
    // (a tile beyond it reads its records from global memory, exactly)
    auto capFor = [&](double m) {
        int v = int(std::min(double(STAGE_MAX), std::ceil(1.25 * m + 64.0)));
        return stageCapLimit > 0 ? std::min(v, stageCapLimit) : v;
    };
    stageCapPlain = capFor(maxStage[0]);
    stageCapAdress = capFor(maxStage[1]);
    stageCapBuild = std::max(stageCapPlain, stageCapAdress);
    mrmd_ctx* c = ctx;
    dreps.reserve(c, int64_t(reps.size()));
    HIPCHECK(hipMemcpyAsync(dreps.p, reps.data(), reps.size() * sizeof(RepD), hipMemcpyHostToDevice, c->stream));
    dtg.reserve(c, ntiles);
    HIPCHECK(hipMemcpyAsync(dtg.p, tg.data(), tg.size() * sizeof(TileG), hipMemcpyHostToDevice, c->stream));
    tiles.reserve(c, ntiles);
    ePart.reserve(c, ntiles);
    hipLaunchKernelGGL(k_bFillD, dim3(gridFor(ntiles)), dim3(BLOCK), 0, c->stream, ePart.p, ntiles, 0.0);
    // a new tiling needs new tile ranges and lists: force a rebuild of every replica
    hipLaunchKernelGGL(k_bFillD, dim3(gridFor(int64_t(reps.size()))), dim3(BLOCK), 0, c->stream, dispAcc.p,
                       int64_t(reps.size()), DBL_MAX);
    checkLaunch();
    sync(c);
    dropGraph();
}

// n steps; a step's post-force integration and the next step's pre-force integration run
// as one pass (k_bPre<true>), the first pre-force and the last post-force pass alone
void mrmd_batch::enqueueSteps(int64_t n)
{
    if (n <= 0) return;
    mrmd_ctx* c = ctx;
    const BS b = state();
    const unsigned ga = gridFor(natoms), gc = gridFor(ncells);
    const unsigned gs = unsigned((ncells * SORT_GROUP + BLOCK - 1) / BLOCK);
    const unsigned nr = unsigned(reps.size());
    hipLaunchKernelGGL(k_bPre<false>, dim3(ga), dim3(BLOCK), 0, c->stream, b);
    for (int64_t k = 0; k < n; ++k)
    {
        hipLaunchKernelGGL(k_bCheck, dim3(gridFor(nr)), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bWrap, dim3(ga), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bZeroCells, dim3(gc), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bCount, dim3(ga), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bScan, dim3(nr), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bFill, dim3(ga), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bSortCells, dim3(gs), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bPermute1, dim3(ga), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bPermute2, dim3(ga), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bTiles, dim3(unsigned((ntiles + BLOCK / WAVE - 1) / (BLOCK / WAVE))), dim3(BLOCK), 0,
                           c->stream, b);
        hipLaunchKernelGGL(k_bGather, dim3(ga), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bBuild, dim3(unsigned(ntiles)), dim3(block), size_t(stageCapBuild + B_PAD) * 24, c->stream,
                           b);
        auto force = [&](auto kernel, int64_t t0, int64_t t1, int capRecs, size_t bytes) {
            if (t1 <= t0) return;
            BS f = b;
            f.tileBase = int32_t(t0);
            f.stageCap = capRecs;
            hipLaunchKernelGGL(kernel, dim3(unsigned(t1 - t0)), dim3(block), size_t(capRecs) * bytes, c->stream, f);
        };
        if (cap_)
        {
            force(k_bForce<true, false>, 0, plainTiles, stageCapPlain, 24);
            force(k_bForce<true, true>, plainTiles, ntiles, stageCapAdress, sizeof(double4));
        }
        else
        {
            force(k_bForce<false, false>, 0, plainTiles, stageCapPlain, 24);
            force(k_bForce<false, true>, plainTiles, ntiles, stageCapAdress, sizeof(double4));
        }
        hipLaunchKernelGGL(k_bCompUpdate, dim3(nr), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bTFUpdate, dim3(nr), dim3(BLOCK), 0, c->stream, b);
        hipLaunchKernelGGL(k_bStepEnd, dim3(1), dim3(64), 0, c->stream, b);
        if (k + 1 < n)
            hipLaunchKernelGGL(k_bPre<true>, dim3(ga), dim3(BLOCK), 0, c->stream, b);
        else
            hipLaunchKernelGGL(k_bPost, dim3(ga), dim3(BLOCK), 0, c->stream, b);
    }
}

// Checkpoint blob of the ensemble's dynamic state: a header of eight int64 (magic, version,
// replicas, atoms, thermodynamic-force bins, compensation bins per replica, 0, 0), then the
// parts of stateParts in their order.  The Verlet lists are not part of it: a save and a
// load both end the current list epoch (every replica rebuilds at the next step), so a run
// that goes on after a save and a run restored from it take bit-identical steps.
namespace
{
constexpr int64_t STATE_MAGIC = 0x3154415453444d4dLL;  // bytes "MMDSTAT1"
constexpr int64_t STATE_VERSION = 2;
constexpr int STATE_HEADER = 8;

struct StatePart
{
    void* dev;
    size_t bytes;
};

std::vector<StatePart> stateParts(mrmd_batch* b)
{
    const size_t na = size_t(b->natoms), nr = b->reps.size(), cb = nr * COMP_BINS;
    std::vector<StatePart> p;
    p.push_back({b->pos.p, na * sizeof(double4)});
    for (int d = 0; d < 3; ++d) p.push_back({b->vel.p + d * b->cap, na * 8});
    for (int d = 0; d < 3; ++d) p.push_back({b->force.p + d * b->cap, na * 8});
    p.push_back({b->id.p, na * 4});
    p.push_back({b->step.p, 8});
    p.push_back({b->rebuilds.p, nr * 8});
    p.push_back({b->tfSamples.p, nr * 8});
    p.push_back({b->tfDensity.p, size_t(b->nbins) * 8});
    p.push_back({b->tfForce.p, size_t(b->nbins) * 8});
    p.push_back({b->comp.p, cb * 8});
    p.push_back({b->compCount.p, cb * 8});
    p.push_back({b->meanComp.p, cb * 8});
    return p;
}

int64_t stateBytes(mrmd_batch* b)
{
    int64_t n = STATE_HEADER * 8;
    for (const StatePart& p : stateParts(b)) n += int64_t(p.bytes);
    return n;
}

// where each atom of the given order sits now
std::vector<int32_t> placeOf(mrmd_batch* b)
{
    const int64_t n = b->natoms;
    std::vector<int32_t> id(static_cast<size_t>(n)), place(static_cast<size_t>(n));
    HIPCHECK(hipMemcpyAsync(id.data(), b->id.p, size_t(n) * 4, hipMemcpyDeviceToHost, b->ctx->stream));
    sync(b->ctx);
    for (int64_t i = 0; i < n; ++i) place[size_t(id[size_t(i)])] = int32_t(i);
    return place;
}

void endListEpoch(mrmd_batch* b)
{
    hipLaunchKernelGGL(k_bFillD, dim3(gridFor(int64_t(b->reps.size()))), dim3(BLOCK), 0, b->ctx->stream,
                       b->dispAcc.p, int64_t(b->reps.size()), DBL_MAX);
    checkLaunch();
}
}  // namespace

extern "C" {

int mrmd_batch_create(mrmd_ctx* c, const mrmd_replica* reps, int64_t nrep, double rc, double skin, mrmd_batch** out)
{
    MRMD_API_BEGIN(c)
    REQUIRE(out && reps && nrep > 0, "invalid arguments");
    REQUIRE(rc > 0.0 && skin >= 0.0, "invalid cutoff / skin");
    const double rn = rc + skin;
    auto* b = new mrmd_batch();
    b->ctx = c;
    b->rc = rc;
    b->skin = skin;
    try
    {
        b->host.assign(reps, reps + nrep);
        int64_t a0 = 0, c0 = 0, h0 = 0;
        for (int64_t r = 0; r < nrep; ++r)
        {
            const mrmd_replica& H = reps[r];
            REQUIRE(H.num_atoms > 0, "replica %lld has no atoms", (long long)r);
            REQUIRE(H.dt > 0.0 && H.mass > 0.0 && H.temperature >= 0.0 && H.gamma >= 0.0,
                    "replica %lld: invalid integration parameters", (long long)r);
            RepD R{};
            R.a0 = a0;
            R.n = H.num_atoms;
            R.c0 = int32_t(c0);
            for (int d = 0; d < 3; ++d)
            {
                REQUIRE(H.box[d] >= 2.0 * rn, "replica %lld: box edge %g is below 2 (rc + skin)", (long long)r,
                        H.box[d]);
                R.nc[d] = int(std::floor(H.box[d] / rn));
                R.L[d] = H.box[d];
                R.dxc[d] = H.box[d] / R.nc[d];
                R.rdxc[d] = 1.0 / R.dxc[d];
            }
            R.dt = H.dt;
            R.zeta = H.gamma;
            R.temperature = H.temperature;
            R.mass = H.mass;
            {
                // the expressions of the per-atom form; the host evaluates them to the same bits
                const double dtm = R.dt / R.mass;
                R.damping = detExp(-R.zeta * dtm);
                R.sigma = sqrt(R.temperature / R.mass * (1.0 - detExp(-2.0 * R.zeta * dtm)));
            }
            R.halfSkin = 0.5 * skin;
            R.seed = H.seed;
            R.adress = H.adress ? 1 : 0;
            R.tfEvery = H.tf_every;
            R.cmp0 = int32_t(r * COMP_BINS);
            if (R.adress)
            {
                REQUIRE(H.tf_bin_width > 0.0 && H.tf_target_density > 0.0 && H.tf_sigma > 0.0,
                        "replica %lld: invalid thermodynamic-force parameters", (long long)r);
                R.slab.center[0] = 0.5 * H.box[0];
                R.slab.center[1] = 0.5 * H.box[1];
                R.slab.center[2] = 0.5 * H.box[2];
                R.slab.atomistic_diameter = H.atomistic_diameter;
                R.slab.hybrid_diameter = H.hybrid_diameter;
                R.slab.nu = H.nu;
                R.slab.interface_type = 0;
                // ThermodynamicForce ctor (ThermodynamicForce.cpp:25-57)
                R.nbins = int32_t(std::ceil(H.box[0] / H.tf_bin_width));
                REQUIRE(R.nbins <= MAX_BINS, "replica %lld: more than %d thermodynamic-force bins", (long long)r,
                        MAX_BINS);
                R.h0 = int32_t(h0);
                R.binWidth = H.box[0] / R.nbins;
                R.invBinWidth = 1.0 / R.binWidth;
                R.binVolume = H.box[1] * H.box[2] * R.binWidth;
                R.forceFactor = H.tf_modulation / H.tf_target_density;
                R.tfSigma = H.tf_sigma;
                R.tfDelta = int32_t(H.tf_range * H.tf_sigma / R.binWidth);
                R.tfSym = H.tf_symmetric;
                R.tfPeriodic = H.tf_periodic;
                h0 += R.nbins;
            }
            a0 += R.n;
            c0 += int64_t(R.nc[0]) * R.nc[1] * R.nc[2];
            b->reps.push_back(R);
        }
        REQUIRE(a0 < (int64_t(1) << 31) - 1 && c0 < (int64_t(1) << 31) - 1, "ensemble too large");
        b->natoms = a0;
        b->cap = std::max<int64_t>(64, (a0 + 63) / 64 * 64);
        b->ncells = c0;
        b->nbins = std::max<int64_t>(h0, 1);
        // list width from the densest replica: Poisson tail of the count within r_n
        double rhoMax = 0.0;
        for (auto& R : b->reps) rhoMax = std::max(rhoMax, double(R.n) / (R.L[0] * R.L[1] * R.L[2]));
        const double mean = 4.0 / 3.0 * PI * rn * rn * rn * rhoMax;
        b->capList = int32_t((int64_t(mean + 8.0 * std::sqrt(mean) + 48.0) + CHUNK - 1) / CHUNK * CHUNK);
        mrmd_ctx* x = c;
        const int64_t na = b->natoms, cp = b->cap, nr = nrep;
        b->pos.reserve(x, cp);
        b->sortedPos.reserve(x, cp);
        b->vel.reserve(x, 3 * cp);
        b->force.reserve(x, 3 * cp);
        b->slotLam.reserve(x, cp);
        b->slotGrad.reserve(x, cp);
        b->atomRep.reserve(x, cp);
        b->cellOf.reserve(x, cp);
        b->sorted.reserve(x, cp);
        b->id.reserve(x, cp);
        b->tmpA.reserve(x, cp);
        b->tmpB.reserve(x, cp);
        b->tmpVel.reserve(x, 3 * cp);
        b->cnt.reserve(x, cp);
        b->cellRep.reserve(x, b->ncells);
        b->counts.reserve(x, b->ncells);
        b->offsets.reserve(x, b->ncells);
        b->fill.reserve(x, b->ncells);
        b->nbr.reserve(x, int64_t(b->capList) * cp);
        b->flag.reserve(x, nr);
        b->err.reserve(x, 1);
        b->dispStep.reserve(x, nr);
        b->dispAcc.reserve(x, nr);
        b->rebuilds.reserve(x, nr);
        b->tfSamples.reserve(x, nr);
        b->energies.reserve(x, 2 * nr);
        b->tfDensity.reserve(x, b->nbins);
        b->tfForce.reserve(x, b->nbins);
        b->comp.reserve(x, nr * COMP_BINS);
        b->compCount.reserve(x, nr * COMP_BINS);
        b->meanComp.reserve(x, nr * COMP_BINS);
        b->step.reserve(x, 1);
        std::vector<int32_t> ar(static_cast<size_t>(na)), cr(static_cast<size_t>(b->ncells));
        for (int64_t r = 0; r < nr; ++r)
        {
            const RepD& R = b->reps[size_t(r)];
            std::fill(ar.begin() + R.a0, ar.begin() + R.a0 + R.n, int32_t(r));
            const int64_t nc = int64_t(R.nc[0]) * R.nc[1] * R.nc[2];
            std::fill(cr.begin() + R.c0, cr.begin() + R.c0 + nc, int32_t(r));
        }
        HIPCHECK(hipMemcpyAsync(b->atomRep.p, ar.data(), size_t(na) * 4, hipMemcpyHostToDevice, x->stream));
        HIPCHECK(hipMemcpyAsync(b->cellRep.p, cr.data(), size_t(b->ncells) * 4, hipMemcpyHostToDevice, x->stream));
        HIPCHECK(hipMemsetAsync(b->vel.p, 0, size_t(3 * cp) * 8, x->stream));
        HIPCHECK(hipMemsetAsync(b->force.p, 0, size_t(3 * cp) * 8, x->stream));
        HIPCHECK(hipMemsetAsync(b->pos.p, 0, size_t(cp) * sizeof(double4), x->stream));
        HIPCHECK(hipMemsetAsync(b->dispStep.p, 0, size_t(nr) * 8, x->stream));
        HIPCHECK(hipMemsetAsync(b->rebuilds.p, 0, size_t(nr) * 8, x->stream));
        HIPCHECK(hipMemsetAsync(b->tfSamples.p, 0, size_t(nr) * 8, x->stream));
        HIPCHECK(hipMemsetAsync(b->flag.p, 0, size_t(nr) * 4, x->stream));
        HIPCHECK(hipMemsetAsync(b->err.p, 0, 4, x->stream));
        HIPCHECK(hipMemsetAsync(b->step.p, 0, 8, x->stream));
        HIPCHECK(hipMemsetAsync(b->tfDensity.p, 0, size_t(b->nbins) * 8, x->stream));
        HIPCHECK(hipMemsetAsync(b->tfForce.p, 0, size_t(b->nbins) * 8, x->stream));
        HIPCHECK(hipMemsetAsync(b->comp.p, 0, size_t(nr) * COMP_BINS * 8, x->stream));
        HIPCHECK(hipMemsetAsync(b->compCount.p, 0, size_t(nr) * COMP_BINS * 8, x->stream));
        HIPCHECK(hipMemsetAsync(b->meanComp.p, 0, size_t(nr) * COMP_BINS * 8, x->stream));
        HIPCHECK(hipMemsetAsync(b->cnt.p, 0, size_t(cp) * 4, x->stream));
        hipLaunchKernelGGL(k_bIota, dim3(gridFor(na)), dim3(BLOCK), 0, x->stream, b->id.p, na);
        checkLaunch();
        // default table: epsilon = sigma = 1, shifted, uncapped
        const double one = 1.0, zero = 0.0;
        std::vector<Pre> t;
        buildTable(t, &zero, &rc, &one, &one, 1, true);
        b->lj = t[0];
        b->layoutTiles();
    }
    catch (...)
    {
        delete b;
        throw;
    }
    *out = b;
    MRMD_API_END
}

int mrmd_batch_destroy(mrmd_batch* b)
{
    if (!b) return MRMD_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
    return MRMD_OK;
}

int mrmd_batch_set_lj(mrmd_batch* b, double cap, double sigma, double epsilon, int shifted)
{
    MRMD_API_BEGIN(b ? b->ctx : nullptr)
    REQUIRE(b && cap >= 0.0 && sigma > 0.0 && epsilon >= 0.0, "invalid LJ parameters");
    std::vector<Pre> t;
    buildTable(t, &cap, &b->rc, &sigma, &epsilon, 1, shifted != 0);
    b->lj = t[0];
    b->cap_ = cap > 0.0;
    b->dropGraph();
    MRMD_API_END
}

int mrmd_batch_set_tiling(mrmd_batch* b, int target, int block)
{
    MRMD_API_BEGIN(b ? b->ctx : nullptr)
    REQUIRE(b && target >= 64 && target <= STAGE_MAX, "target staged records must lie in [64, %d]", STAGE_MAX);
    REQUIRE(block == 64 || block == 128 || block == 192 || block == 256, "workgroup size must be 64, 128, 192 or 256");
    b->targetStaged = target;
    b->block = block;
    b->layoutTiles();
    MRMD_API_END
}

int mrmd_batch_set_stage_cap(mrmd_batch* b, int limit)
{
    MRMD_API_BEGIN(b ? b->ctx : nullptr)
    REQUIRE(b && limit >= 0 && limit <= STAGE_MAX, "the stage limit must lie in [0, %d]", STAGE_MAX);
    b->stageCapLimit = limit;
    b->layoutTiles();
    MRMD_API_END
}

int mrmd_batch_get_tiling(const mrmd_batch* b, int* target, int* block, int32_t* tileCells)
{
    if (!b) return MRMD_ERR_ARG;
    if (target) *target = b->targetStaged;
    if (block) *block = b->block;
    if (tileCells)
        for (size_t r = 0; r < b->reps.size(); ++r) tileCells[r] = b->reps[r].K;
    return MRMD_OK;
}

// host arrays are in the order the atoms were given; on the device an atom sits in its
// replica's slot order of the last rebuild (place)
int mrmd_batch_upload(mrmd_batch* b, int field, const double* host)
{
    MRMD_API_BEGIN(b ? b->ctx : nullptr)
    REQUIRE(b && host && (field == MRMD_POS || field == MRMD_VEL), "upload takes POS or VEL");
    const int64_t n = b->natoms;
    mrmd_ctx* c = b->ctx;
    const std::vector<int32_t> place = placeOf(b);
    if (field == MRMD_POS)
    {
        std::vector<double4> h(static_cast<size_t>(n));
        for (int64_t i = 0; i < n; ++i)
            h[size_t(place[size_t(i)])] = make_double4(host[3 * i], host[3 * i + 1], host[3 * i + 2], 0.0);
        HIPCHECK(hipMemcpyAsync(b->pos.p, h.data(), size_t(n) * sizeof(double4), hipMemcpyHostToDevice, c->stream));
        sync(c);
        // positions changed outside the time loop: every replica rebuilds at the next step
        hipLaunchKernelGGL(k_bFillD, dim3(gridFor(int64_t(b->reps.size()))), dim3(BLOCK), 0, c->stream, b->dispAcc.p,
                           int64_t(b->reps.size()), DBL_MAX);
        checkLaunch();
    }
    else
    {
        std::vector<double> h(static_cast<size_t>(3 * b->cap), 0.0);
        for (int64_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) h[size_t(d * b->cap + place[size_t(i)])] = host[3 * i + d];
        HIPCHECK(hipMemcpyAsync(b->vel.p, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->stream));
    }
    sync(c);
    MRMD_API_END
}

int mrmd_batch_download(mrmd_batch* b, int field, double* host)
{
    MRMD_API_BEGIN(b ? b->ctx : nullptr)
    REQUIRE(b && host && (field == MRMD_POS || field == MRMD_VEL || field == MRMD_FORCE),
            "download takes POS, VEL or FORCE");
    const int64_t n = b->natoms;
    mrmd_ctx* c = b->ctx;
    const std::vector<int32_t> place = placeOf(b);
    if (field == MRMD_POS)
    {
        std::vector<double4> h(static_cast<size_t>(n));
        HIPCHECK(hipMemcpyAsync(h.data(), b->pos.p, size_t(n) * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        for (int64_t i = 0; i < n; ++i)
        {
            const double4& q = h[size_t(place[size_t(i)])];
            host[3 * i] = q.x;
            host[3 * i + 1] = q.y;
            host[3 * i + 2] = q.z;
        }
    }
    else
    {
        std::vector<double> h(static_cast<size_t>(3 * b->cap));
        HIPCHECK(hipMemcpyAsync(h.data(), field == MRMD_VEL ? b->vel.p : b->force.p, h.size() * 8,
                                hipMemcpyDeviceToHost, c->stream));
        sync(c);
        for (int64_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) host[3 * i + d] = h[size_t(d * b->cap + place[size_t(i)])];
    }
    MRMD_API_END
}

int mrmd_batch_run(mrmd_batch* b, int64_t nsteps, int graphSteps)
{
    MRMD_API_BEGIN(b ? b->ctx : nullptr)
    REQUIRE(b && nsteps >= 0 && graphSteps >= 0, "invalid arguments");
    mrmd_ctx* c = b->ctx;
    KTimer t(c, "batch_run");
    int64_t done = 0;
    if (graphSteps > 0 && nsteps >= graphSteps)
    {
        if (!b->exec || b->graphSteps != graphSteps)
        {
            b->dropGraph();
            HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            b->enqueueSteps(graphSteps);
            HIPCHECK(hipStreamEndCapture(c->stream, &b->graph));
            HIPCHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graphSteps = graphSteps;
        }
        for (; done + graphSteps <= nsteps; done += graphSteps) HIPCHECK(hipGraphLaunch(b->exec, c->stream));
    }
    if (done < nsteps) b->enqueueSteps(nsteps - done);
    checkLaunch();
    int32_t e = 0;
    HIPCHECK(hipMemcpyAsync(&e, b->err.p, 4, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    b->steps += nsteps;
    if (e & 1) fail(MRMD_ERR_STATE, "a neighbour row exceeded the list width %d (density too high)", b->capList);
    MRMD_API_END
}

int mrmd_batch_energies(mrmd_batch* b, double* potential, double* kinetic)
{
    MRMD_API_BEGIN(b ? b->ctx : nullptr)
    REQUIRE(b, "null batch");
    mrmd_ctx* c = b->ctx;
    const size_t nr = b->reps.size();
    hipLaunchKernelGGL(k_bEnergies, dim3(unsigned(nr)), dim3(BLOCK), 0, c->stream, b->state(), b->energies.p);
    checkLaunch();
    std::vector<double> h(2 * nr);
    HIPCHECK(hipMemcpyAsync(h.data(), b->energies.p, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    for (size_t r = 0; r < nr; ++r)
    {
        if (potential) potential[r] = h[2 * r];
        if (kinetic) kinetic[r] = h[2 * r + 1];
    }
    MRMD_API_END
}

int mrmd_batch_stats(mrmd_batch* b, int64_t* rebuilds, int64_t* steps)
{
    MRMD_API_BEGIN(b ? b->ctx : nullptr)
    REQUIRE(b, "null batch");
    mrmd_ctx* c = b->ctx;
    if (rebuilds)
        HIPCHECK(hipMemcpyAsync(rebuilds, b->rebuilds.p, b->reps.size() * 8, hipMemcpyDeviceToHost, c->stream));
    if (steps) HIPCHECK(hipMemcpyAsync(steps, b->step.p, 8, hipMemcpyDeviceToHost, c->stream));
    sync(c);
    MRMD_API_END
}

int mrmd_batch_state_bytes(const mrmd_batch* b, int64_t* bytes)
{
    if (!b || !bytes) return MRMD_ERR_ARG;
    *bytes = stateBytes(const_cast<mrmd_batch*>(b));
    return MRMD_OK;
}

int mrmd_batch_state_save(mrmd_batch* b, void* host, int64_t bytes)
{
    MRMD_API_BEGIN(b ? b->ctx : nullptr)
    REQUIRE(b && host, "null argument");
    const int64_t need = stateBytes(b);
    REQUIRE(bytes >= need, "the buffer holds %lld of %lld bytes", (long long)bytes, (long long)need);
    mrmd_ctx* c = b->ctx;
    const int64_t hdr[STATE_HEADER] = {STATE_MAGIC, STATE_VERSION, int64_t(b->reps.size()), b->natoms, b->nbins,
                                       COMP_BINS, 0, 0};
    char* out = static_cast<char*>(host);
    std::memcpy(out, hdr, sizeof hdr);
    size_t off = sizeof hdr;
    for (const StatePart& p : stateParts(b))
    {
        HIPCHECK(hipMemcpyAsync(out + off, p.dev, p.bytes, hipMemcpyDeviceToHost, c->stream));
        off += p.bytes;
    }
    sync(c);
    endListEpoch(b);
    sync(c);
    MRMD_API_END
}

int mrmd_batch_state_load(mrmd_batch* b, const void* host, int64_t bytes)
{
    MRMD_API_BEGIN(b ? b->ctx : nullptr)
    REQUIRE(b && host, "null argument");
    const int64_t need = stateBytes(b);
    REQUIRE(bytes == need, "the state holds %lld bytes, this ensemble's %lld", (long long)bytes, (long long)need);
    int64_t hdr[STATE_HEADER];
    std::memcpy(hdr, host, sizeof hdr);
    REQUIRE(hdr[0] == STATE_MAGIC && hdr[1] == STATE_VERSION, "not an ensemble state (version %d)",
            int(STATE_VERSION));
    REQUIRE(hdr[2] == int64_t(b->reps.size()) && hdr[3] == b->natoms && hdr[4] == b->nbins && hdr[5] == COMP_BINS,
            "the state is of %lld replicas / %lld atoms / %lld bins, this ensemble %lld / %lld / %lld",
            (long long)hdr[2], (long long)hdr[3], (long long)hdr[4], (long long)b->reps.size(),
            (long long)b->natoms, (long long)b->nbins);
    mrmd_ctx* c = b->ctx;
    const char* in = static_cast<const char*>(host);
    size_t off = sizeof hdr;
    for (const StatePart& p : stateParts(b))
    {
        HIPCHECK(hipMemcpyAsync(p.dev, in + off, p.bytes, hipMemcpyHostToDevice, c->stream));
        off += p.bytes;
    }
    sync(c);
    endListEpoch(b);
    sync(c);
    MRMD_API_END
}

int mrmd_batch_tf_get(mrmd_batch* b, int64_t replica, int which, double* host, int64_t n)
{
    MRMD_API_BEGIN(b ? b->ctx : nullptr)
    REQUIRE(b && host && replica >= 0 && replica < int64_t(b->reps.size()), "invalid arguments");
    const RepD& R = b->reps[size_t(replica)];
    REQUIRE(R.adress, "replica %lld is not an AdResS replica", (long long)replica);
    REQUIRE(n >= R.nbins, "the output holds %lld of %d bins", (long long)n, R.nbins);
    const double* src = (which == 0 ? b->tfForce.p : b->tfDensity.p) + R.h0;
    HIPCHECK(hipMemcpyAsync(host, src, size_t(R.nbins) * 8, hipMemcpyDeviceToHost, b->ctx->stream));
    sync(b->ctx);
    MRMD_API_END
}

}  // extern "C"
