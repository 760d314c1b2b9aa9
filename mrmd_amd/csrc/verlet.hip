// Verlet list build (Cabana VerletList<HalfNeighborTag, VerletLayout2D> semantics) in a
// tiled, LDS-staged form.
//
// All records (local and ghost) are binned on Cabana's grid and sorted by cell.  A tile
// is a run of K cells along z in a x a neighbouring (i, j) columns; its rows are one
// contiguous range of sorted slots per column, and every cell of its stencils lies in the
// (a + 2R)^2 columns around it, each again a contiguous slot range.  One workgroup per tile
// copies those ranges into LDS once and every row scans its stencil cells (Cabana's i, j, k
// order and min-distance prune) from LDS.  An entry is the 15-bit LDS slot of the neighbour plus a flag bit for
// the pairs Cabana's position-lexicographic half discriminator keeps, so the exact half
// list is a sub-list and the force kernels stage the same tile and never gather from
// global memory in their inner loop.  Global-index rows (for the API, the adjoint and the
// half-list kernels) are converted from this form on demand.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>

#include <type_traits>

#include "nlist.hpp"

using namespace mrmd_hip;

namespace
{
struct TileGeom
{
    int n0, n1, n2, K, tilesPerCol, R, cols, nbj;
    int o0, o1, o2;  ///< first cell of the region that holds rows, per axis
    int e0, e1, e2;  ///< its extent in cells
};

// first column (i0, j0) and z-run [k0, k1) of tile t; tiles cover only the cells that hold
// list rows (a ghost layer of the grid holds none), so no lane of a tile sits on a cell
// of ghost records
__device__ inline void tileCoords(const TileGeom& tg, int64_t t, int& i0, int& j0, int& k0, int& k1)
{
    const int col = int(t / tg.tilesPerCol);
    const int kb = int(t % tg.tilesPerCol);
    i0 = tg.o0 + (col / tg.nbj) * tg.cols;
    j0 = tg.o1 + (col % tg.nbj) * tg.cols;
    k0 = tg.o2 + kb * tg.K;
    k1 = min(k0 + tg.K, tg.o2 + tg.e2);
}

// tile of a cell inside the row region
__device__ inline int64_t tileOfCell(const TileGeom& tg, int cell)
{
    const int k = cell % tg.n2, ij = cell / tg.n2, i = ij / tg.n1, j = ij % tg.n1;
    return (int64_t((i - tg.o0) / tg.cols) * tg.nbj + (j - tg.o1) / tg.cols) * tg.tilesPerCol + (k - tg.o2) / tg.K;
}

// one thread per tile: row ranges, neighbourhood column ranges and their LDS bases
// the maximum of v over the 64 lanes of the wave (every lane active)
__device__ inline int waveMaxInt(int v)
{
    for (int off = WAVE / 2; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, WAVE));
    return v;
}

__device__ void tileSetup(const TileGeom& tg, int64_t t, const int32_t* offsets, const int32_t* counts,
                          TileInfo* tiles, int& stagedOut);

// one lane per tile; the largest staged neighbourhood goes to *maxStaged with one atomic per
// wave
__global__ __launch_bounds__(BLOCK) void k_tileSetup(TileGeom tg, int64_t ntiles, const int32_t* offsets, const int32_t* counts,
                            TileInfo* tiles, int32_t* maxStaged)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    int staged = 0;
    if (t < ntiles) tileSetup(tg, t, offsets, counts, tiles, staged);
    staged = waveMaxInt(staged);
    if (threadIdx.x % WAVE == 0 && staged > 0) atomicMax(maxStaged, staged);
}

__device__ void tileSetup(const TileGeom& tg, int64_t t, const int32_t* offsets, const int32_t* counts,
                          TileInfo* tiles, int& stagedOut)
{
    int i0, j0, k0, k1;
    tileCoords(tg, t, i0, j0, k0, k1);
    TileInfo& ti = tiles[t];
    const int span = tg.cols + 2 * tg.R;
    const int kl = max(k0 - tg.R, 0), kh = min(k1 + tg.R, tg.n2);
    const int iEnd = min(i0 + tg.cols, tg.o0 + tg.e0), jEnd = min(j0 + tg.cols, tg.o1 + tg.e1);
    // every range's table entries are loaded first (independent loads in flight together),
    // then the ranges are laid out
    int first[MAX_RANGES], last[MAX_RANGES], lastCount[MAX_RANGES], selfFirst[MAX_RANGES], selfLast[MAX_RANGES],
        selfCount[MAX_RANGES];
#pragma unroll
    for (int q = 0; q < MAX_RANGES; ++q)
    {
        const int a = q / 5, b = q % 5;  // (a, b) of a span x span neighbourhood, span <= 5
        const int ii = i0 - tg.R + a, jj = j0 - tg.R + b;
        const bool in = a < span && b < span && ii >= 0 && ii < tg.n0 && jj >= 0 && jj < tg.n1;
        const bool self = in && ii >= i0 && ii < iEnd && jj >= j0 && jj < jEnd;
        const int col = (ii * tg.n1 + jj) * tg.n2;
        first[q] = in ? offsets[col + kl] : 0;
        last[q] = in ? offsets[col + kh - 1] : 0;
        lastCount[q] = in ? counts[col + kh - 1] : 0;
        selfFirst[q] = self ? offsets[col + k0] : 0;
        selfLast[q] = self ? offsets[col + k1 - 1] : 0;
        selfCount[q] = self ? counts[col + k1 - 1] : -1;
    }
    int nr = 0, staged = 0, nself = 0, nrows = 0;
#pragma unroll
    for (int q = 0; q < MAX_RANGES; ++q)
    {
        const int a = q / 5, b = q % 5;
        if (a >= span || b >= span) continue;
        const int ii = i0 - tg.R + a, jj = j0 - tg.R + b;
        ti.base[nr] = staged;
        ti.begin[nr] = first[q];
        if (ii >= 0 && ii < tg.n0 && jj >= 0 && jj < tg.n1)
        {
            staged += last[q] + lastCount[q] - first[q];
            if (selfCount[q] >= 0)
            {
                ti.selfRange[nself] = nr;
                ti.selfBegin[nself] = selfFirst[q];
                ti.selfPrefix[nself] = nrows;
                nrows += selfLast[q] + selfCount[q] - selfFirst[q];
                ++nself;
            }
        }
        ++nr;
    }
    ti.selfPrefix[nself] = nrows;
    ti.nr = nr;
    ti.staged = staged;
    ti.nself = nself;
    ti.nrows = nrows;
    stagedOut = staged;
}

// bounding box of the cells of the list rows (records [rowLo, rowHi) of the binning):
// box[0..2] minimum cell index per axis, box[3..5] maximum.  A grid-stride pass with a
// wave and then a workgroup reduction, so that only a few hundred workgroups post atomics
// to the six words.
constexpr int BOX_BLOCKS = 512;
// q = c / n and r = c % n for 0 <= c < 2^31, 1 <= n < 2^31 (rn = 1.0 / n): the quotient from
// a double product, off by at most one, corrected by the remainder's range
__device__ inline int divCell(int c, int n, double rn, int& r)
{
    int q = int(double(c) * rn);
    r = c - q * n;
    if (r < 0)
    {
        --q;
        r += n;
    }
    else if (r >= n)
    {
        ++q;
        r -= n;
    }
    return q;
}
__global__ __launch_bounds__(BLOCK) void k_rowCellBox(const int32_t* cellOf, int64_t rowLo, int64_t rowHi, int n1,
                                                       int n2, int32_t* box)
{
    __shared__ int red[6][BLOCK / WAVE];
    int lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {-1, -1, -1};
    const double r1 = 1.0 / double(n1), r2 = 1.0 / double(n2);
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    // eight independent loads per lane in flight, then their cells
    constexpr int U = 8;
    for (int64_t k0 = rowLo + int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k0 < rowHi; k0 += U * stride)
    {
        int c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = k0 + u * stride < rowHi ? cellOf[k0 + u * stride] : -1;
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            if (c[u] < 0) continue;
            int kc, jc;
            const int ij = divCell(c[u], n2, r2, kc);
            const int ic = divCell(ij, n1, r1, jc);
            const int v[3] = {ic, jc, kc};
            for (int d = 0; d < 3; ++d)
            {
                lo[d] = min(lo[d], v[d]);
                hi[d] = max(hi[d], v[d]);
            }
        }
    }
    for (int d = 0; d < 3; ++d)
        for (int off = WAVE / 2; off > 0; off >>= 1)
        {
            lo[d] = min(lo[d], __shfl_xor(lo[d], off, WAVE));
            hi[d] = max(hi[d], __shfl_xor(hi[d], off, WAVE));
        }
    const int w = threadIdx.x / WAVE;
    if (threadIdx.x % WAVE == 0)
        for (int d = 0; d < 3; ++d)
        {
            red[d][w] = lo[d];
            red[3 + d][w] = hi[d];
        }
    __syncthreads();
    if (threadIdx.x < 6)
    {
        const int d = threadIdx.x;
        int v = red[d][0];
        for (int q = 1; q < int(blockDim.x) / WAVE; ++q) v = d < 3 ? min(v, red[d][q]) : max(v, red[d][q]);
        if (d < 3 ? v != INT32_MAX : v >= 0)
        {
            if (d < 3)
                atomicMin(&box[d], v);
            else
                atomicMax(&box[d], v);
        }
    }
}

__global__ void k_initBox(int32_t* box)
{
    if (threadIdx.x < 6) box[threadIdx.x] = threadIdx.x < 3 ? INT32_MAX : -1;
}

constexpr int MAX_SPAN = 64;  // cells of a tile's column range along z (K + 2R)
// records of LDS past the staged neighbourhood that a candidate pass may read (its group of
// four runs up to three records beyond the column run; their bits are cleared)
constexpr int STAGE_PAD = 4;

// One lane per row.  Every row scans its stencil columns in Cabana's i, j, k order with the
// min-distance prune (the surviving cells of a column as one slot range, cut by binary
// search to the z window of the row, the records being z-sorted), 32 candidates per pass in
// groups of four (independent LDS reads and distances, hit bits shifted in by a funnel shift),
// then the hits appended in slot order.
// WIDE: list offsets in 64-bit arithmetic (lists of 2^31 entries or more); 32-bit otherwise
// XCD: tiles in the chip's XCD-contiguous order (neighbouring tiles, which share most of
// their staged records, on one XCD and its L2); otherwise plain block order
template <int R, bool WIDE, bool XCD, bool GS>
__global__ __launch_bounds__(BLOCK) void k_verletTiled(const double4* __restrict__ sortedPos,
                                                       const int32_t* __restrict__ sorted,
                                                       const int32_t* __restrict__ cellOf,
                                                       const int32_t* __restrict__ offsets,
                                                       const TileInfo* __restrict__ tiles, TileGeom tg, GridD g,
                                                       int64_t rowLo, int64_t rowHi, double rsq, uint16_t* nbr,
                                                       int64_t strideS, int cap, int32_t* cnt, int32_t* maxCnt,
                                                       int stagedCap, double4* gstage, int64_t gstageRecs,
                                                       double bandLo, double bandHi, int8_t* tileSkip)
{
    constexpr int S = 2 * R + 1;
    extern __shared__ double4 stRaw[];
    __shared__ TileInfo ti;
    const int64_t t = XCD ? int64_t(xcdBlock()) : int64_t(blockIdx.x);
    // the staged {x, y, z} records: in LDS, or (GS, dense inputs) in this tile's slice of
    // global memory, with the per-cell start table alone in LDS
    double* const st = GS ? reinterpret_cast<double*>(gstage + t * gstageRecs) : reinterpret_cast<double*>(stRaw);
    // after the staged records and their pad: the LDS slot of the first record of cell
    // kl + z of range r, cstart[r * cz + z] with cz = K + 2R + 1; entry z = nk closes the
    // range (ranges outside the grid are empty).  Sized to the tile height, so that short
    // tiles leave LDS for more workgroups.
    const int cz = tg.K + 2 * R + 1;
    int* cstart = GS ? reinterpret_cast<int*>(stRaw)
                     : reinterpret_cast<int*>(reinterpret_cast<double*>(stRaw) + 3 * (int64_t(stagedCap) + STAGE_PAD));
    if (threadIdx.x == 0) ti = tiles[t];
    __syncthreads();
    if (ti.nrows == 0) return;
    int i0, j0, k0, k1;
    tileCoords(tg, t, i0, j0, k0, k1);
    if (bandLo <= bandHi)
    {
        // x range of the neighbourhood's cells (the edge cells take every record beyond
        // the grid; widened by 1e-9 of a cell for records binned across a face by rounding)
        const int ia = max(i0 - R, 0), ib = min(i0 + tg.cols - 1 + R, tg.n0 - 1);
        const double xlo = ia == 0 ? -HUGE_VAL : g.min[0] + double(ia) * g.dx[0] - 1e-9 * g.dx[0];
        const double xhi = ib == tg.n0 - 1 ? HUGE_VAL : g.min[0] + double(ib + 1) * g.dx[0] + 1e-9 * g.dx[0];
        const bool skip = xhi < bandLo || xlo > bandHi;
        if (threadIdx.x == 0) tileSkip[t] = skip ? 1 : 0;
        if (skip)
        {
            for (int row = threadIdx.x; row < ti.nrows; row += blockDim.x)
            {
                int self;
                cnt[tileRow(ti, row, self)] = 0;
            }
            return;
        }
    }
    const int span = tg.cols + 2 * R;
    const int kl = max(k0 - R, 0), nk = min(k1 + R, tg.n2) - kl;
    // the first row's dependent loads (slot -> record -> cell) are issued before the staging
    int s = -1, self = 0, cell = 0;
    int32_t p = -1;
    if (int(threadIdx.x) < ti.nrows)
    {
        s = tileRow(ti, threadIdx.x, self);
        p = sorted[s];
        cell = cellOf[p];
    }
    stageTile<true>(ti, sortedPos, st);
    for (int x = threadIdx.x; x < ti.nr * (nk + 1); x += blockDim.x)
    {
        const int r = x / (nk + 1), z = x % (nk + 1);
        const int ii = i0 - R + r / span, jj = j0 - R + r % span;
        int v = ti.base[r];
        if (ii >= 0 && ii < tg.n0 && jj >= 0 && jj < tg.n1)
        {
            if (z < nk)
                v = ti.base[r] + (offsets[(ii * tg.n1 + jj) * tg.n2 + kl + z] - ti.begin[r]);
            else
                v = r + 1 < ti.nr ? ti.base[r + 1] : ti.staged;
        }
        cstart[r * cz + z] = v;
    }
    __syncthreads();
    using Off = std::conditional_t<WIDE, int64_t, uint32_t>;
    int localMax = 0;
    for (int row = threadIdx.x; row < ti.nrows; row += blockDim.x)
    {
        if (row != int(threadIdx.x))
        {
            s = tileRow(ti, row, self);
            p = sorted[s];
            cell = cellOf[p];
        }
        if (p < rowLo || p >= rowHi)
        {
            cnt[s] = 0;
            continue;
        }
        const int qSelf = ldsSlot(ti, ti.selfRange[self], s);
        const double xpx = st[3 * qSelf], xpy = st[3 * qSelf + 1], xpz = st[3 * qSelf + 2];
        const int ck = cell % tg.n2, i = (cell / tg.n2) / tg.n1, j = (cell / tg.n2) % tg.n1;
        // per-axis terms of Cabana's minDistanceToPoint (r per axis of the cell box, summed
        // as r2 = ((0 + rx^2) + ry^2) + rz^2, so lateral part + rz^2 rounds exactly as the
        // full sum) and of the z window (the lateral distances shrunk by 1e-9 of a cell,
        // far more than a record binned across a face by rounding can stray)
        double rx2[S], ry2[S], wx2[S], wy2[S], rz2[S];
#pragma unroll
        for (int a = 0; a < S; ++a)
        {
            const double cx = g.min[0] + (double(i - R + a) + 0.5) * g.dx[0];
            const double cy = g.min[1] + (double(j - R + a) + 0.5) * g.dx[1];
            const double czc = g.min[2] + (double(ck - R + a) + 0.5) * g.dx[2];
            double r = fabs(xpx - cx) - 0.5 * g.dx[0];
            r = r > 0.0 ? r : 0.0;
            rx2[a] = r * r;
            r = fabs(xpy - cy) - 0.5 * g.dx[1];
            r = r > 0.0 ? r : 0.0;
            ry2[a] = r * r;
            r = fabs(xpz - czc) - 0.5 * g.dx[2];
            r = r > 0.0 ? r : 0.0;
            rz2[a] = r * r;
            double e = fabs(xpx - cx) - 0.5 * g.dx[0] - 1e-9 * g.dx[0];
            e = e > 0.0 ? e : 0.0;
            wx2[a] = e * e;
            e = fabs(xpy - cy) - 0.5 * g.dx[1] - 1e-9 * g.dx[1];
            e = e > 0.0 ? e : 0.0;
            wy2[a] = e * e;
        }
        // entries are collected eight at a time in four registers (each append shifts the
        // eight 16-bit places down by one and puts the entry in the top place, so the first
        // entry ends in the lowest) and stored as one 16-byte chunk: chunk k of slot s is
        // chunk k * strideS + s of the list
        uint4* const chunkBase = reinterpret_cast<uint4*>(nbr) + s;
        uint32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0;
        int c = 0;
        auto append = [&](uint16_t v) {
            b0 = __builtin_amdgcn_alignbit(b1, b0, 16u);
            b1 = __builtin_amdgcn_alignbit(b2, b1, 16u);
            b2 = __builtin_amdgcn_alignbit(b3, b2, 16u);
            b3 = __builtin_amdgcn_alignbit(uint32_t(v), b3, 16u);
            if ((c & 7) == 7) chunkBase[Off(c >> 3) * Off(strideS)] = make_uint4(b0, b1, b2, b3);
            ++c;
        };
#pragma unroll
        for (int da = -R; da <= R; ++da)
        {
            const int ii = i + da;
#pragma unroll
            for (int db = -R; db <= R; ++db)
            {
                const int jj = j + db;
                int lo = 0, hi = 0;
                if (ii >= 0 && ii < tg.n0 && jj >= 0 && jj < tg.n1)
                {
                    const int* cs0 = cstart + ((ii - (i0 - R)) * span + (jj - (j0 - R))) * cz;
                    // Cabana's prune per cell of the column: the distance is unimodal in k, so
                    // the surviving cells are one run [kLo, kHi]
                    const double lat = rx2[da + R] + ry2[db + R];
                    int kLo = INT_MAX, kHi = -1;
#pragma unroll
                    for (int dc = 0; dc < S; ++dc)
                    {
                        const int kk = ck - R + dc;
                        if (kk >= 0 && kk < tg.n2 && lat + rz2[dc] <= rsq)
                        {
                            kLo = min(kLo, kk);
                            kHi = kk;
                        }
                    }
                    if (kLo <= kHi)
                    {
                        // records inside a column run are z-sorted: the candidates that can pass
                        // are those whose (dz)^2, as the test computes it, stays within rsq less
                        // the squared lateral distance from the row to the column (widened by
                        // 1e-12 rsq, more than the rounding of the test's sums), so the cut
                        // changes no decision; the window's lower edge lies in cell kLo and its
                        // upper edge in cell kHi, so each binary search runs over one cell
                        const double zsq = rsq - (wx2[da + R] + wy2[db + R]) + 1e-12 * rsq;
                        int a = cs0[kLo - kl], b = cs0[kLo + 1 - kl];
                        while (a < b)
                        {
                            const int m = (a + b) >> 1;
                            const double dz = xpz - st[3 * m + 2];
                            if (dz > 0.0 && dz * dz > zsq) a = m + 1;
                            else b = m;
                        }
                        lo = a;
                        a = max(lo, cs0[kHi - kl]);
                        b = cs0[kHi + 1 - kl];
                        while (a < b)
                        {
                            const int m = (a + b) >> 1;
                            const double dz = xpz - st[3 * m + 2];
                            if (dz < 0.0 && dz * dz > zsq) b = m;
                            else a = m + 1;
                        }
                        hi = a;
                    }
                }
                for (int qb = lo; qb < hi; qb += 32)
                {
                    // branch-free pass over up to 32 candidates in groups of four; each
                    // group's reads may run up to three records past the run (inside the LDS
                    // pad), their bits are cleared below.  hits (and, for the columns of the
                    // row's own x index, the sign bits of dx and the dx == 0 ties) take one bit
                    // per candidate, shifted in from the right (first candidate highest)
                    // A bit is shifted in by one funnel shift of the sign word of a double:
                    // rsq - d^2 is negative exactly when d^2 > rsq (a rounded difference keeps
                    // the sign of the exact one and is +0 only for equal operands), and dx
                    // is negative exactly when its sign bit is set (xp - xn is never -0)
                    uint32_t miss = 0, neg = 0, tie = 0;
                    const int nq = min(32, hi - qb);
                    const double* q = st + 3 * qb;
                    int m = 0;
                    for (; m < nq; m += 4)
                    {
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                        {
                            const double dx = xpx - q[3 * (m + u)];
                            const double dy = xpy - q[3 * (m + u) + 1];
                            const double dz = xpz - q[3 * (m + u) + 2];
                            const double r2 = dx * dx + dy * dy + dz * dz;
                            miss = __builtin_amdgcn_alignbit(miss, uint32_t(__double2hiint(rsq - r2)), 31u);
                            if (da == 0)
                            {
                                unsigned carry;
                                neg = __builtin_amdgcn_alignbit(neg, uint32_t(__double2hiint(dx)), 31u);
                                tie = __builtin_addc(tie, tie, unsigned(dx == 0.0), &carry);
                            }
                        }
                    }
                    // align: candidate u at bit 31 - u; drop the candidates past the run and
                    // the row itself
                    const int sh = 32 - m;  // m is 4..32
                    uint32_t hits = ~miss << sh;
                    hits &= ~((1u << (32 - nq)) - 1u);
                    if (qSelf >= qb && qSelf < qb + nq) hits &= ~(0x80000000u >> (qSelf - qb));
                    if (da == 0)
                    {
                        neg <<= sh;
                        tie <<= sh;
                    }
                    // append the hits in slot order with their half-list bit, Cabana's
                    // lexicographic discriminator (xn > xp || (xn == xp && (yn > yp || (yn ==
                    // yp && zn > zp)))): columns of a larger x index hold larger x (binning is
                    // monotone in x), smaller index smaller x; in the row's own x index the
                    // sign of dx decides, and a tie dx == 0 compares y and z
                    auto appendHits = [&](auto checked) {
                        while (hits)
                        {
                            const int u = __builtin_clz(hits);
                            hits ^= 0x80000000u >> u;
                            uint32_t half = da > 0 ? 1u : 0u;
                            if (da == 0)
                            {
                                half = (neg << u) >> 31;
                                if ((tie << u) & 0x80000000u)
                                {
                                    const double ny = q[3 * u + 1], nz = q[3 * u + 2];
                                    half = uint32_t((ny > xpy) | ((ny == xpy) & (nz > xpz)));
                                }
                            }
                            const uint16_t v = uint16_t((qb + u) | (half << 15));
                            if (!decltype(checked)::value || c < cap)
                                append(v);
                            else
                                ++c;
                        }
                    };
                    // a row that overflows the capacity counts its further hits without storing
                    if (c + __builtin_popcount(hits) <= cap)
                        appendHits(std::false_type{});
                    else
                        appendHits(std::true_type{});
                }
            }
        }
        // pad the last chunk with the sentinel slot (a record beyond every cutoff), so the
        // force kernels' chunk loop needs no bounds test
        const int width = c;
        while ((c & 7) != 0 && c < cap) append(uint16_t(sentinelSlot(stagedCap)));
        cnt[s] = min(width, cap);  // the true width goes to maxCnt; reads stay inside the list
        localMax = max(localMax, width);
    }
    // one atomic per wave (every lane of the workgroup reaches this point)
    localMax = waveMaxInt(localMax);
    if (threadIdx.x % WAVE == 0 && localMax > 0) atomicMax(maxCnt, localMax);
}

// tiled rows (per sorted slot, LDS slots) -> global rows (per record begin+r, record indices)
__global__ void k_toGlobal(const uint16_t* nbr16, const int32_t* cnt16, int64_t strideS, const TileInfo* tiles,
                           TileGeom tg, const int32_t* sorted, const int32_t* cellOf, int64_t nslots, int64_t rowLo,
                           int64_t rowHi, uint32_t* nbr, int32_t* cnt, int64_t stride)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const int32_t p = sorted[s];
    if (p < rowLo || p >= rowHi) return;
    const int64_t r = p - rowLo;
    const TileInfo& ti = tiles[tileOfCell(tg, cellOf[p])];
    const int c = cnt16[s];
    for (int k = 0; k < c; ++k)
    {
        const uint16_t e = nbr16[entryIndex(k, s, strideS)];
        const int q = e & 0x7fff;
        int rr = 0;
        while (rr + 1 < ti.nr && ti.base[rr + 1] <= q) ++rr;
        const int32_t jrec = sorted[ti.begin[rr] + (q - ti.base[rr])];
        nbr[int64_t(k) * stride + r] = uint32_t(jrec) | ((e & 0x8000) ? 0x80000000u : 0u);
    }
    cnt[r] = c;
}

__global__ void k_gatherSorted(const double4* pos, const int32_t* sorted, int64_t n, double4* out)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s < n) out[s] = pos[sorted[s]];
}

__global__ void k_scatterSorted(const double4* sortedPos, const int32_t* sorted, int64_t n, double4* pos)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) pos[sorted[k]] = sortedPos[k];
}

__global__ void k_invertSorted(const int32_t* sorted, int64_t n, int32_t* slotOf)
{
    const int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s < n) slotOf[sorted[s]] = int32_t(s);
}

__global__ void k_zeroI32v(int32_t* p, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

__global__ void k_halfExtract(const uint32_t* nbr, const int32_t* cnt, int64_t nrows, int64_t stride, int32_t* hnbr,
                              int32_t* hcnt)
{
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int c = cnt[r];
    int h = 0;
    for (int k = 0; k < c; ++k)
    {
        const uint32_t e = nbr[int64_t(k) * stride + r];
        if (e & 0x80000000u) hnbr[int64_t(h++) * stride + r] = int32_t(e & 0x7fffffffu);
    }
    hcnt[r] = h;
}

__global__ void k_countWithin(const double4* pos, const uint32_t* nbr, const int32_t* cnt, int64_t begin, int64_t rows,
                              int64_t stride, double r2, unsigned long long* out)
{
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned long long h = 0, f = 0;
    if (r < rows)
    {
        const double4 pi = pos[begin + r];
        for (int k = 0; k < cnt[r]; ++k)
        {
            const uint32_t e = nbr[int64_t(k) * stride + r];
            const double4 pj = pos[e & 0x7fffffffu];
            const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            if (dx * dx + dy * dy + dz * dz <= r2)
            {
                ++f;
                if (e & 0x80000000u) ++h;
            }
        }
    }
    for (int off = WAVE / 2; off > 0; off >>= 1)
    {
        h += __shfl_xor(h, off, WAVE);
        f += __shfl_xor(f, off, WAVE);
    }
    if (threadIdx.x % WAVE == 0)
    {
        atomicAdd(&out[0], h);
        atomicAdd(&out[1], f);
    }
}

__global__ void k_rowsToHost(const uint32_t* nbr, const int32_t* cnt, int64_t nrows, int64_t stride, int32_t* out,
                             int64_t rowStride, int mask)
{
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int c = cnt[r];
    for (int k = 0; k < c && k < rowStride; ++k)
    {
        uint32_t e = nbr[int64_t(k) * stride + r];
        if (mask) e &= 0x7fffffffu;
        out[r * rowStride + k] = int32_t(e);
    }
}

TileGeom geom(const mrmd_verlet_list* vl)
{
    const GridD& g = vl->bin.grid;
    const int a = vl->tileCols;
    return TileGeom{g.n[0], g.n[1], g.n[2], vl->tileK, vl->tilesPerCol, vl->range, a,
                    (vl->rowExtent[1] + a - 1) / a,
                    vl->rowOrigin[0], vl->rowOrigin[1], vl->rowOrigin[2],
                    vl->rowExtent[0], vl->rowExtent[1], vl->rowExtent[2]};
}

// cell box of the list rows, from the binning (one device reduction and a readback)
void rowRegion(mrmd_verlet_list* vl)
{
    mrmd_ctx* ctx = vl->ctx;
    const int64_t rows = vl->end - vl->begin;
    for (int d = 0; d < 3; ++d)
    {
        vl->rowOrigin[d] = 0;
        vl->rowExtent[d] = 0;
    }
    if (rows <= 0) return;
    int32_t* box = static_cast<int32_t*>(ctxScratch(ctx, 2, 64)) + 8;
    hipLaunchKernelGGL(k_initBox, dim3(1), dim3(64), 0, ctx->stream, box);
    hipLaunchKernelGGL(k_rowCellBox, dim3(std::min<unsigned>(gridFor(rows), BOX_BLOCKS)), dim3(BLOCK), 0, ctx->stream,
                       vl->bin.cellOf.p, vl->begin,
                       vl->end, vl->bin.grid.n[1], vl->bin.grid.n[2], box);
    checkLaunch();
    int32_t h[6];
    readbackI32(ctx, box, 6, h);
    for (int d = 0; d < 3; ++d)
    {
        // rows exist, so every axis holds at least one cell of them
        if (h[d] < 0 || h[3 + d] < h[d] || h[3 + d] >= vl->bin.grid.n[d])
            fail(MRMD_ERR_STATE, "cell box of the list rows out of the grid: axis %d cells [%d, %d] of %d (%lld rows)",
                 d, h[d], h[3 + d], vl->bin.grid.n[d], (long long)rows);
        vl->rowOrigin[d] = h[d];
        vl->rowExtent[d] = h[3 + d] - h[d] + 1;
    }
}

void setupTiles(mrmd_verlet_list* vl, int K)
{
    mrmd_ctx* ctx = vl->ctx;
    vl->tileK = K;
    vl->tilesPerCol = (vl->rowExtent[2] + K - 1) / K;
    const int a = vl->tileCols;
    vl->ntiles = int64_t((vl->rowExtent[0] + a - 1) / a) * ((vl->rowExtent[1] + a - 1) / a) * vl->tilesPerCol;
    vl->tiles.ensure(ctx, std::max<int64_t>(vl->ntiles, 1));
    int32_t* dmax = static_cast<int32_t*>(ctxScratch(ctx, 2, 64));
    hipLaunchKernelGGL(k_zeroI32v, dim3(1), dim3(64), 0, ctx->stream, dmax, int64_t(1));
    if (vl->ntiles > 0)
        hipLaunchKernelGGL(k_tileSetup, dim3(gridFor(vl->ntiles)), dim3(BLOCK), 0, ctx->stream, geom(vl), vl->ntiles,
                           vl->bin.offsets.p, vl->bin.counts.p, vl->tiles.p, dmax);
    checkLaunch();
    int32_t h[1];
    readbackI32(ctx, dmax, 1, h);
    vl->maxStaged = h[0];
}

// records whose atom count is not 1 (molecule lists only)
__global__ void k_countNotOne(const int64_t* numAtoms, int64_t n, int32_t* out)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n && numAtoms[k] != 1) atomicAdd(out, 1);
}

// numAtoms (molecule lists) is checked for records of one atom; returns whether all are
bool verletBuild(mrmd_verlet_list* vl, const double4* pos, int64_t nall, int64_t begin, int64_t end, double rn,
                 double ratio, const double* gmin, const double* gmax, int64_t maxNeigh,
                 const int64_t* numAtoms = nullptr)
{
    mrmd_ctx* ctx = vl->ctx;
    REQUIRE(begin >= 0 && begin <= end && end <= nall, "invalid row range [%lld, %lld) of %lld", (long long)begin,
            (long long)end, (long long)nall);
    REQUIRE(rn > 0.0 && ratio > 0.0, "invalid radius / cell ratio");
    REQUIRE(nall < (int64_t(1) << 31) - 1, "more than 2^31 records");
    KTimer t(ctx, "verlet_build");
    vl->nrows = nall;
    vl->begin = begin;
    vl->end = end;
    vl->rn = rn;
    vl->ratio = ratio;
    vl->range = int(std::ceil(1.0 / ratio));
    if (vl->range > 2) fail(MRMD_ERR_UNSUP, "cell ratio %g needs a stencil wider than 5 cells", ratio);
    const double gs = ratio * rn;
    const double delta[3] = {gs, gs, gs};
    const GridD g = makeGrid(gmin, gmax, delta);
    vl->sentinel = std::max({gmax[0], gmax[1], gmax[2]}) + SENTINEL_GAP + 4.0 * rn;
    // the binning's z sort also writes each record's slot (slotOf)
    vl->slotOf.ensure(ctx, std::max<int64_t>(nall, 1));
    vl->bin.build(ctx, pos, 0, nall, g, true, true, nullptr, nullptr, vl->slotOf.p);
    rowRegion(vl);
    vl->globalValid = false;
    vl->halfValid = false;
    vl->banded = vl->bandLo <= vl->bandHi;
    vl->slotOfValid = true;
    vl->sortedFresh = false;
    // tile shape: a x a columns (a <= 3; 1 when the stencil reaches two cells), height K
    // chosen to stage ~1650 records; shrink until every tile fits in LDS
    const int R = vl->range;
    int a = std::max(1, std::min(vl->tileCols, R == 1 ? 3 : 1));
    const int W = a + 2 * R;
    int K = vl->tileK > 0 ? std::min(vl->tileK, g.n[2]) : 0;
    if (K <= 0)
    {
        const double regionCells = double(std::max(vl->rowExtent[0], 1)) * std::max(vl->rowExtent[1], 1) *
                                   std::max(vl->rowExtent[2], 1);
        const double perCell = double(std::max<int64_t>(end - begin, 1)) / regionCells;
        K = int(1650.0 / (W * W * std::max(perCell, 1e-3))) - 2 * R;
        K = std::max(1, std::min(K, g.n[2]));
    }
    K = std::min(K, MAX_SPAN - 2 * R);
    vl->globalStage = false;
    for (;;)
    {
        vl->tileCols = a;
        setupTiles(vl, K);
        if (vl->maxStaged <= vl->stageLimit) break;
        if (K > 1)
            K = std::max(1, K / 2);
        else if (a > 1)
            --a;
        else
        {
            // dense input: even one cell column by one cell exceeds the LDS limit, so each
            // tile stages its neighbourhood in its own slice of global memory
            vl->globalStage = true;
            break;
        }
    }
    if (vl->globalStage)
    {
        REQUIRE(vl->maxStaged < 32767, "a tile neighbourhood of %d records exceeds the 15-bit entry slots",
                vl->maxStaged);
        vl->gstageRecs = (int64_t(vl->maxStaged) + STAGE_PAD + 1 + 3) / 4 * 4;
        vl->gstage.ensure(ctx, std::max<int64_t>(vl->ntiles, 1) * vl->gstageRecs);
    }
    if (vl->banded) vl->tileSkip.ensure(ctx, std::max<int64_t>(vl->ntiles, 1));
    vl->strideS = std::max<int64_t>(64, (nall + 63) / 64 * 64);
    vl->cnt16.ensure(ctx, vl->strideS);
    hipLaunchKernelGGL(k_zeroI32v, dim3(gridFor(nall)), dim3(BLOCK), 0, ctx->stream, vl->cnt16.p, nall);
    if (vl->capFull <= 0) vl->capFull = int32_t(std::max<int64_t>(2 * maxNeigh + 16, 96));
    vl->capFull = (vl->capFull + CHUNK - 1) / CHUNK * CHUNK;
    int32_t* dmax = static_cast<int32_t*>(ctxScratch(ctx, 2, 64)) + 1;
    int32_t* dmulti = dmax + 1;
    hipLaunchKernelGGL(k_zeroI32v, dim3(1), dim3(64), 0, ctx->stream, dmulti, int64_t(1));
    if (numAtoms && nall > 0)
        hipLaunchKernelGGL(k_countNotOne, dim3(gridFor(nall)), dim3(BLOCK), 0, ctx->stream, numAtoms, nall, dmulti);
    bool mono = false;
    // staged records, then the per-cell start table of (a + 2R)^2 ranges x (K + 2R + 1)
    const int stagedCap = std::max(vl->maxStaged, 1);
    const size_t lds = (vl->globalStage ? 0 : size_t(stagedCap + STAGE_PAD) * Staged<true>::bytes) +
                       size_t((vl->tileCols + 2 * R) * (vl->tileCols + 2 * R) * (vl->tileK + 2 * R + 1)) * sizeof(int);
    for (int attempt = 0; attempt < 2; ++attempt)
    {
        vl->nbr16.ensure(ctx, int64_t(vl->capFull) * vl->strideS);
        hipLaunchKernelGGL(k_zeroI32v, dim3(1), dim3(64), 0, ctx->stream, dmax, int64_t(1));
        if (vl->ntiles > 0 && end > begin)
        {
            KTimer tk(ctx, "verlet_tiled");
            auto launch = [&](auto kernel) {
                hipLaunchKernelGGL(kernel, dim3(unsigned(vl->ntiles)), dim3(vl->block), lds, ctx->stream,
                                   vl->bin.sortedPos.p, vl->bin.sorted.p, vl->bin.cellOf.p, vl->bin.offsets.p,
                                   vl->tiles.p, geom(vl), g, begin, end, rn * rn, vl->nbr16.p, vl->strideS,
                                   vl->capFull, vl->cnt16.p, dmax, stagedCap, vl->gstage.p, vl->gstageRecs,
                                   vl->bandLo, vl->bandHi, vl->tileSkip.p);
            };
            const bool wide = vl->forceWide || int64_t(vl->capFull) * vl->strideS >= (int64_t(1) << 31);
            // plain block order: measured faster here than the XCD-contiguous order (1.54 against
            // 1.59 ms at 4M atoms), each XCD's L2 then serves a spread of the grid's columns
            if (vl->globalStage)
            {
                if (R == 1)
                    wide ? launch(k_verletTiled<1, true, false, true>) : launch(k_verletTiled<1, false, false, true>);
                else
                    wide ? launch(k_verletTiled<2, true, false, true>) : launch(k_verletTiled<2, false, false, true>);
            }
            else if (R == 1)
                wide ? launch(k_verletTiled<1, true, false, false>) : launch(k_verletTiled<1, false, false, false>);
            else
                wide ? launch(k_verletTiled<2, true, false, false>) : launch(k_verletTiled<2, false, false, false>);
        }
        checkLaunch();
        int32_t hp[2];
        readbackI32(ctx, dmax, 2, hp);
        const int32_t hmax = hp[0];
        mono = hp[1] == 0;
        vl->maxFull = hmax;
        if (hmax <= vl->capFull) break;
        // Cabana VerletLayout2D: re-count and refill with the larger width
        vl->capFull = (hmax + hmax / 8 + 8 + CHUNK - 1) / CHUNK * CHUNK;
    }
    if (vl->maxFull > vl->capFull) fail(MRMD_ERR_STATE, "neighbour rows did not fit after a re-count");
    return mono;
}
}  // namespace

void mrmd_verlet_list::refreshSorted(const double4* pos)
{
    if (nrows > 0)
        hipLaunchKernelGGL(k_gatherSorted, dim3(gridFor(nrows)), dim3(BLOCK), 0, ctx->stream, pos, bin.sorted.p, nrows,
                           bin.sortedPos.p);
    checkLaunch();
}

void mrmd_verlet_list::positionsFromSlots(double4* pos)
{
    if (nrows > 0)
        hipLaunchKernelGGL(k_scatterSorted, dim3(gridFor(nrows)), dim3(BLOCK), 0, ctx->stream, bin.sortedPos.p,
                           bin.sorted.p, nrows, pos);
    checkLaunch();
}

// (the error paths of the time loop: a launch error is left for the caller's next call)
void mrmd_verlet_list::positionsFromSlotsNoThrow(double4* pos) noexcept
{
    if (nrows > 0)
        hipLaunchKernelGGL(k_scatterSorted, dim3(gridFor(nrows)), dim3(BLOCK), 0, ctx->stream, bin.sortedPos.p,
                           bin.sorted.p, nrows, pos);
}

const int32_t* mrmd_verlet_list::ensureSlotOf()
{
    if (!slotOfValid)
    {
        slotOf.ensure(ctx, std::max<int64_t>(nrows, 1));
        if (nrows > 0)
            hipLaunchKernelGGL(k_invertSorted, dim3(gridFor(nrows)), dim3(BLOCK), 0, ctx->stream, bin.sorted.p, nrows,
                               slotOf.p);
        checkLaunch();
        slotOfValid = true;
    }
    return slotOf.p;
}

void mrmd_verlet_list::ensureGlobal()
{
    if (banded)
        fail(MRMD_ERR_STATE, "the list holds only the rows within reach of the slab's weighted region (built "
                             "inside the AdResS time loop); rebuild it for this use");
    if (globalValid) return;
    const int64_t rows = end - begin;
    stride = std::max<int64_t>(64, (rows + 63) / 64 * 64);
    cnt.ensure(ctx, stride);
    nbr.ensure(ctx, int64_t(std::max(maxFull, 1)) * stride);
    if (rows > 0)
    {
        hipLaunchKernelGGL(k_zeroI32v, dim3(gridFor(rows)), dim3(BLOCK), 0, ctx->stream, cnt.p, rows);
        hipLaunchKernelGGL(k_toGlobal, dim3(gridFor(nrows)), dim3(BLOCK), 0, ctx->stream, nbr16.p, cnt16.p, strideS,
                           tiles.p, geom(this), bin.sorted.p, bin.cellOf.p, nrows, begin, end, nbr.p, cnt.p, stride);
    }
    checkLaunch();
    globalValid = true;
}

void mrmd_verlet_list::ensureHalf()
{
    if (halfValid) return;
    ensureGlobal();
    const int64_t rows = end - begin;
    capHalf = std::max(maxFull, 1);
    hnbr.ensure(ctx, int64_t(capHalf) * stride);
    hcnt.ensure(ctx, stride);
    if (rows > 0)
        hipLaunchKernelGGL(k_halfExtract, dim3(gridFor(rows)), dim3(BLOCK), 0, ctx->stream, nbr.p, cnt.p, rows, stride,
                           hnbr.p, hcnt.p);
    checkLaunch();
    halfValid = true;
}

extern "C" {

int mrmd_verlet_create(mrmd_ctx* c, mrmd_verlet_list** out)
{
    MRMD_API_BEGIN(c)
    REQUIRE(out, "null output");
    auto* v = new mrmd_verlet_list();
    v->ctx = c;
    *out = v;
    MRMD_API_END
}

int mrmd_verlet_destroy(mrmd_verlet_list* v)
{
    if (!v) return MRMD_OK;
    (void)hipStreamSynchronize(v->ctx->stream);
    delete v;
    return MRMD_OK;
}

int mrmd_verlet_set_tiling(mrmd_verlet_list* vl, int tile_cells, int tile_columns, int block)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl && tile_cells >= 0 && tile_columns >= 1 && tile_columns <= 3, "invalid tiling");
    REQUIRE(block == 64 || block == 128 || block == 192 || block == 256, "workgroup size must be 64, 128, 192 or 256");
    vl->tileK = tile_cells;
    vl->tileCols = tile_columns;
    vl->block = block;
    MRMD_API_END
}

int mrmd_verlet_set_stage_limit(mrmd_verlet_list* vl, int limit)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl && limit >= 0 && limit <= STAGE_MAX, "the stage limit must lie in [0, %d]", STAGE_MAX);
    vl->stageLimit = limit > 0 ? limit : STAGE_MAX;
    MRMD_API_END
}

int mrmd_verlet_set_wide_offsets(mrmd_verlet_list* vl, int on)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl, "null list");
    vl->forceWide = on != 0;
    MRMD_API_END
}

int mrmd_verlet_get_tiling(const mrmd_verlet_list* vl, int* tile_cells, int* tile_columns, int* block,
                           int* max_staged)
{
    if (!vl) return MRMD_ERR_ARG;
    if (tile_cells) *tile_cells = vl->tileK;
    if (tile_columns) *tile_columns = vl->tileCols;
    if (block) *block = vl->block;
    if (max_staged) *max_staged = vl->maxStaged;
    return MRMD_OK;
}

int mrmd_verlet_build(mrmd_verlet_list* vl, const mrmd_atoms* a, int64_t begin, int64_t end, double rn, double ratio,
                      const double* gmin, const double* gmax, int64_t maxNeigh)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl && a && gmin && gmax, "invalid arguments");
    vl->monoOf = nullptr;
    verletBuild(vl, a->pos.p, a->nlocal + a->nghost, begin, end, rn, ratio, gmin, gmax, maxNeigh);
    MRMD_API_END
}

int mrmd_verlet_build_molecules(mrmd_verlet_list* vl, const mrmd_molecules* m, int64_t begin, int64_t end, double rn,
                                double ratio, const double* gmin, const double* gmax, int64_t maxNeigh)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl && m && gmin && gmax, "invalid arguments");
    vl->monoOf = nullptr;
    if (verletBuild(vl, m->pos.p, m->nlocal + m->nghost, begin, end, rn, ratio, gmin, gmax, maxNeigh, m->numAtoms.p))
    {
        vl->monoOf = m;
        vl->monoVersion = m->layoutVersion;
    }
    MRMD_API_END
}

int mrmd_verlet_half_counts(mrmd_verlet_list* vl, int32_t* counts, int64_t n, int64_t* total)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl, "null list");
    vl->ensureHalf();
    const int64_t rows = vl->end - vl->begin;
    std::vector<int32_t> h(static_cast<size_t>(rows));
    if (rows) HIPCHECK(hipMemcpyAsync(h.data(), vl->hcnt.p, size_t(rows) * 4, hipMemcpyDeviceToHost, vl->ctx->stream));
    sync(vl->ctx);
    int64_t tot = 0;
    for (auto v : h) tot += v;
    if (total) *total = tot;
    if (counts)
        for (int64_t i = 0; i < n; ++i)
            counts[i] = (i >= vl->begin && i < vl->end) ? h[size_t(i - vl->begin)] : 0;
    MRMD_API_END
}

int mrmd_verlet_half_neighbors(mrmd_verlet_list* vl, int32_t* host, int64_t n, int64_t rowStride)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl && host && rowStride > 0, "invalid arguments");
    vl->ensureHalf();
    mrmd_ctx* ctx = vl->ctx;
    const int64_t rows = std::min<int64_t>(n, vl->end - vl->begin);
    if (rows <= 0) return MRMD_OK;
    DBuf<int32_t> dev;
    dev.reserve(ctx, rows * rowStride);
    hipLaunchKernelGGL(k_rowsToHost, dim3(gridFor(rows)), dim3(BLOCK), 0, ctx->stream,
                       reinterpret_cast<const uint32_t*>(vl->hnbr.p), vl->hcnt.p, rows, vl->stride, dev.p, rowStride, 0);
    HIPCHECK(hipMemcpyAsync(host, dev.p, size_t(rows * rowStride) * 4, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    MRMD_API_END
}

int mrmd_verlet_full_neighbors(mrmd_verlet_list* vl, int32_t* host, int32_t* counts, int64_t n, int64_t rowStride)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl && rowStride > 0, "invalid arguments");
    vl->ensureGlobal();
    mrmd_ctx* ctx = vl->ctx;
    const int64_t rows = std::min<int64_t>(n, vl->end - vl->begin);
    if (rows <= 0) return MRMD_OK;
    if (counts) HIPCHECK(hipMemcpyAsync(counts, vl->cnt.p, size_t(rows) * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (host)
    {
        DBuf<int32_t> dev;
        dev.reserve(ctx, rows * rowStride);
        hipLaunchKernelGGL(k_rowsToHost, dim3(gridFor(rows)), dim3(BLOCK), 0, ctx->stream, vl->nbr.p, vl->cnt.p, rows,
                           vl->stride, dev.p, rowStride, 1);
        HIPCHECK(hipMemcpyAsync(host, dev.p, size_t(rows * rowStride) * 4, hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
    }
    sync(ctx);
    MRMD_API_END
}

int mrmd_verlet_count_within(mrmd_verlet_list* vl, const mrmd_atoms* a, double r, int64_t* half, int64_t* full)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl && a && vl->nrows <= a->size, "invalid arguments");
    vl->ensureGlobal();
    mrmd_ctx* ctx = vl->ctx;
    const int64_t rows = vl->end - vl->begin;
    unsigned long long* d = static_cast<unsigned long long*>(ctxScratch(ctx, 2, 64));
    HIPCHECK(hipMemsetAsync(d, 0, 16, ctx->stream));
    if (rows > 0)
        hipLaunchKernelGGL(k_countWithin, dim3(gridFor(rows)), dim3(BLOCK), 0, ctx->stream, a->pos.p, vl->nbr.p,
                           vl->cnt.p, vl->begin, rows, vl->stride, r * r, d);
    checkLaunch();
    unsigned long long h[2];
    HIPCHECK(hipMemcpyAsync(h, d, 16, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    if (half) *half = int64_t(h[0]);
    if (full) *full = int64_t(h[1]);
    MRMD_API_END
}

int mrmd_verlet_info(const mrmd_verlet_list* vl, int64_t* rows, int32_t* maxFull, int64_t* totalFull, int32_t* dims)
{
    if (!vl) return MRMD_ERR_ARG;
    MRMD_API_BEGIN(vl->ctx)
    if (rows) *rows = vl->end - vl->begin;
    if (maxFull) *maxFull = vl->maxFull;
    if (totalFull)
    {
        auto* v = const_cast<mrmd_verlet_list*>(vl);
        const int64_t r = v->end - v->begin;
        std::vector<int32_t> h(static_cast<size_t>(r));
        if (r)
        {
            v->ensureGlobal();
            HIPCHECK(hipMemcpyAsync(h.data(), v->cnt.p, size_t(r) * 4, hipMemcpyDeviceToHost, v->ctx->stream));
            sync(v->ctx);
        }
        int64_t t = 0;
        for (auto c : h) t += c;
        *totalFull = t;
    }
    if (dims)
        for (int d = 0; d < 3; ++d) dims[d] = vl->bin.grid.n[d];
    MRMD_API_END
}

}  // extern "C"
