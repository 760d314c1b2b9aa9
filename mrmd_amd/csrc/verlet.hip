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
#include <cmath>

#include <type_traits>

#include "nlist.hpp"

using namespace mrmd_hip;

namespace
{
struct TileGeom
{
    int n0, n1, n2, K, tilesPerCol, R, cols, nbj;
};

// first column (i0, j0) and z-run [k0, k1) of tile t
__device__ inline void tileCoords(const TileGeom& tg, int64_t t, int& i0, int& j0, int& k0, int& k1)
{
    const int col = int(t / tg.tilesPerCol);
    const int kb = int(t % tg.tilesPerCol);
    i0 = (col / tg.nbj) * tg.cols;
    j0 = (col % tg.nbj) * tg.cols;
    k0 = kb * tg.K;
    k1 = min(k0 + tg.K, tg.n2);
}

__device__ inline int64_t tileOfCell(const TileGeom& tg, int cell)
{
    const int k = cell % tg.n2, ij = cell / tg.n2, i = ij / tg.n1, j = ij % tg.n1;
    return (int64_t(i / tg.cols) * tg.nbj + j / tg.cols) * tg.tilesPerCol + k / tg.K;
}

// one thread per tile: row ranges, neighbourhood column ranges and their LDS bases
__global__ void k_tileSetup(TileGeom tg, int64_t ntiles, const int32_t* offsets, const int32_t* counts,
                            TileInfo* tiles, int32_t* maxStaged)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    int i0, j0, k0, k1;
    tileCoords(tg, t, i0, j0, k0, k1);
    TileInfo& ti = tiles[t];
    const int span = tg.cols + 2 * tg.R;
    const int kl = max(k0 - tg.R, 0), kh = min(k1 + tg.R, tg.n2);
    int nr = 0, staged = 0, nself = 0, nrows = 0;
    for (int a = 0; a < span; ++a)
        for (int b = 0; b < span; ++b, ++nr)
        {
            const int ii = i0 - tg.R + a, jj = j0 - tg.R + b;
            ti.base[nr] = staged;
            ti.begin[nr] = 0;
            if (ii < 0 || ii >= tg.n0 || jj < 0 || jj >= tg.n1) continue;
            const int c0 = (ii * tg.n1 + jj) * tg.n2 + kl;
            const int c1 = (ii * tg.n1 + jj) * tg.n2 + (kh - 1);
            ti.begin[nr] = offsets[c0];
            staged += offsets[c1] + counts[c1] - offsets[c0];
            if (ii >= i0 && ii < i0 + tg.cols && jj >= j0 && jj < j0 + tg.cols)
            {
                const int r0 = (ii * tg.n1 + jj) * tg.n2 + k0;
                const int r1 = (ii * tg.n1 + jj) * tg.n2 + (k1 - 1);
                ti.selfRange[nself] = nr;
                ti.selfBegin[nself] = offsets[r0];
                ti.selfPrefix[nself] = nrows;
                nrows += offsets[r1] + counts[r1] - offsets[r0];
                ++nself;
            }
        }
    ti.selfPrefix[nself] = nrows;
    ti.nr = nr;
    ti.staged = staged;
    ti.nself = nself;
    ti.nrows = nrows;
    atomicMax(maxStaged, staged);
}

constexpr int MAX_SPAN = 64;  // cells of a tile's column range along z (K + 2R)

// One lane per row.  Every row scans its stencil cells in Cabana's i, j, k order with the
// min-distance prune (the surviving cells of a column as one slot range, cut by binary
// search to the z window of the row, the records being z-sorted), eight candidates per pass
// (independent LDS reads and distances, then the appends in slot order).
// WIDE: list offsets in 64-bit arithmetic (lists of 2^31 entries or more); 32-bit otherwise
template <int R, bool WIDE>
__global__ __launch_bounds__(BLOCK) void k_verletTiled(const double4* __restrict__ sortedPos,
                                                       const int32_t* __restrict__ sorted,
                                                       const int32_t* __restrict__ cellOf,
                                                       const int32_t* __restrict__ offsets,
                                                       const TileInfo* __restrict__ tiles, TileGeom tg, GridD g,
                                                       int64_t rowLo, int64_t rowHi, double rsq, uint16_t* nbr,
                                                       int64_t strideS, int cap, int32_t* cnt, int32_t* maxCnt,
                                                       int stagedCap)
{
    extern __shared__ double4 stRaw[];
    void* st = stRaw;
    __shared__ TileInfo ti;
    // after the staged records: the LDS slot of the first record of cell kl + z of range r,
    // cstart[r * cz + z] with cz = K + 2R + 1; entry z = nk closes the range (ranges outside
    // the grid are empty).  Sized to the tile height, so that short tiles leave LDS for more
    // workgroups.
    const int cz = tg.K + 2 * R + 1;
    int* cstart = reinterpret_cast<int*>(static_cast<double*>(st) + 3 * int64_t(stagedCap));
    const int64_t t = blockIdx.x;
    if (threadIdx.x == 0) ti = tiles[t];
    __syncthreads();
    if (ti.nrows == 0) return;
    int i0, j0, k0, k1;
    tileCoords(tg, t, i0, j0, k0, k1);
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
        // entry k of this row: rowBase[(k / 8) * chunkStride + k % 8] (entryIndex)
        using Off = std::conditional_t<WIDE, int64_t, int32_t>;
        uint16_t* const rowBase = nbr + int64_t(s) * CHUNK;
        const Off chunkStride = Off(strideS) * CHUNK;
        const double4 xp = Staged<true>::get(st, qSelf);
        const int ck = cell % tg.n2, i = (cell / tg.n2) / tg.n1, j = (cell / tg.n2) % tg.n1;
        int c = 0;
        for (int ii = max(i - R, 0); ii < min(i + R + 1, tg.n0); ++ii)
            for (int jj = max(j - R, 0); jj < min(j + R + 1, tg.n1); ++jj)
            {
                const int* cs0 = cstart + ((ii - (i0 - R)) * span + (jj - (j0 - R))) * cz;
                // Cabana's min-distance prune per cell: along z the distance to a cell is
                // unimodal in k (cell centres are monotone, the x and y terms fixed), so the
                // cells that survive form one run [kLo, kHi] and the candidates one slot range
                int kLo = max(ck - R, 0), kHi = min(ck + R + 1, tg.n2) - 1;
                while (kLo <= kHi && g.minDistSqr(xp.x, xp.y, xp.z, ii, jj, kLo) > rsq) ++kLo;
                while (kHi >= kLo && g.minDistSqr(xp.x, xp.y, xp.z, ii, jj, kHi) > rsq) --kHi;
                if (kLo > kHi) continue;
                // records inside a column run are z-sorted: the candidates that can pass are
                // those whose (dz)^2, as the test computes it, stays within rsq less the
                // squared distance from the row to the column's cross-section (every record of
                // the column lies at least that far away laterally); the lateral distance is
                // shrunk by 1e-9 of a cell, far more than a record binned across a face by
                // rounding can stray, and the bound widened by 1e-12 rsq, more than the
                // rounding of the test's sums, so the cut changes no decision
                // the window's lower edge lies in cell kLo and its upper edge in cell kHi (a
                // cell wholly outside the window in z is out of reach and was pruned), so each
                // binary search runs over one cell's records
                double ex = fabs(xp.x - (g.min[0] + (double(ii) + 0.5) * g.dx[0])) - 0.5 * g.dx[0] - 1e-9 * g.dx[0];
                double ey = fabs(xp.y - (g.min[1] + (double(jj) + 0.5) * g.dx[1])) - 0.5 * g.dx[1] - 1e-9 * g.dx[1];
                ex = ex > 0.0 ? ex : 0.0;
                ey = ey > 0.0 ? ey : 0.0;
                const double zsq = rsq - (ex * ex + ey * ey) + 1e-12 * rsq;
                int lo, hi;
                {
                    int a = cs0[kLo - kl], b = cs0[kLo + 1 - kl];  // first slot not below the window
                    while (a < b)
                    {
                        const int m = (a + b) >> 1;
                        const double dz = xp.z - Staged<true>::get(st, m).z;
                        if (dz > 0.0 && dz * dz > zsq) a = m + 1;
                        else b = m;
                    }
                    lo = a;
                    a = max(lo, cs0[kHi - kl]);
                    b = cs0[kHi + 1 - kl];  // first slot above the window
                    while (a < b)
                    {
                        const int m = (a + b) >> 1;
                        const double dz = xp.z - Staged<true>::get(st, m).z;
                        if (dz < 0.0 && dz * dz > zsq) b = m;
                        else a = m + 1;
                    }
                    hi = a;
                }
                const int q1 = hi;
                for (int qb = lo; qb < q1; qb += 32)
                {
                    // branch-free passes over up to 32 candidates: hit bits
                    uint32_t hits = 0;
                    const int nq = min(32, q1 - qb);
                    for (int u0 = 0; u0 < nq; u0 += 4)
                    {
#pragma unroll
                        for (int u = u0; u < u0 + 4; ++u)
                        {
                            // reads stay inside the run; the bits past its end and the row's
                            // own slot are cleared below
                            const double4 n = Staged<true>::get(st, min(qb + u, q1 - 1));
                            const double dx = xp.x - n.x;
                            const double dy = xp.y - n.y;
                            const double dz = xp.z - n.z;
                            hits |= uint32_t(dx * dx + dy * dy + dz * dz <= rsq) << u;
                        }
                    }
                    hits &= nq == 32 ? 0xffffffffu : (1u << nq) - 1u;
                    if (qSelf >= qb && qSelf < qb + nq) hits &= ~(1u << (qSelf - qb));
                    // append the hits in slot order with their half-list bit (Cabana's
                    // lexicographic order, evaluated without branches on all three coordinates)
                    while (hits)
                    {
                        const int u = __builtin_ctz(hits);
                        hits &= hits - 1;
                        const double4 n = Staged<true>::get(st, qb + u);
                        const bool half = (n.x > xp.x) | ((n.x == xp.x) & ((n.y > xp.y) | ((n.y == xp.y) & (n.z > xp.z))));
                        if (c < cap) rowBase[Off(c >> 3) * chunkStride + (c & 7)] = uint16_t((qb + u) | (half ? 0x8000 : 0));
                        ++c;
                    }
                }
            }
        // pad the last chunk with the sentinel slot (a record beyond every cutoff), so the
        // force kernels' chunk loop needs no bounds test
        for (int k = c; k < ((c + CHUNK - 1) / CHUNK) * CHUNK && k < cap; ++k)
            rowBase[Off(k >> 3) * chunkStride + (k & 7)] = uint16_t(sentinelSlot(stagedCap));
        cnt[s] = min(c, cap);  // the true width goes to maxCnt; reads stay inside the list
        localMax = max(localMax, c);
    }
    if (localMax > 0) atomicMax(maxCnt, localMax);
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
    return TileGeom{g.n[0], g.n[1], g.n[2], vl->tileK, vl->tilesPerCol, vl->range, vl->tileCols,
                    (g.n[1] + vl->tileCols - 1) / vl->tileCols};
}

void setupTiles(mrmd_verlet_list* vl, int K)
{
    mrmd_ctx* ctx = vl->ctx;
    const GridD& g = vl->bin.grid;
    vl->tileK = K;
    vl->tilesPerCol = (g.n[2] + K - 1) / K;
    const int a = vl->tileCols;
    vl->ntiles = int64_t((g.n[0] + a - 1) / a) * ((g.n[1] + a - 1) / a) * vl->tilesPerCol;
    vl->tiles.ensure(ctx, std::max<int64_t>(vl->ntiles, 1));
    int32_t* dmax = static_cast<int32_t*>(ctxScratch(ctx, 2, 64));
    hipLaunchKernelGGL(k_zeroI32v, dim3(1), dim3(64), 0, ctx->stream, dmax, int64_t(1));
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
    vl->bin.build(ctx, pos, 0, nall, g, true, true);
    vl->globalValid = false;
    vl->halfValid = false;
    vl->slotOfValid = false;
    vl->sortedFresh = false;
    // tile shape: a x a columns (a <= 3; 1 when the stencil reaches two cells), height K
    // chosen to stage ~1650 records; shrink until every tile fits in LDS
    const int R = vl->range;
    int a = std::max(1, std::min(vl->tileCols, R == 1 ? 3 : 1));
    const int W = a + 2 * R;
    int K = vl->tileK > 0 ? std::min(vl->tileK, g.n[2]) : 0;
    if (K <= 0)
    {
        const double perCell = double(std::max<int64_t>(nall, 1)) / double(g.total());
        K = int(1650.0 / (W * W * std::max(perCell, 1e-3))) - 2 * R;
        K = std::max(1, std::min(K, g.n[2]));
    }
    K = std::min(K, MAX_SPAN - 2 * R);
    for (;;)
    {
        vl->tileCols = a;
        setupTiles(vl, K);
        if (vl->maxStaged <= STAGE_MAX) break;
        if (K > 1)
            K = std::max(1, K / 2);
        else if (a > 1)
            --a;
        else
            fail(MRMD_ERR_UNSUP, "a tile neighbourhood holds %d records (limit %d): density too high for the tiled list",
                 vl->maxStaged, STAGE_MAX);
    }
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
    const size_t lds = size_t(stagedCap) * Staged<true>::bytes +
                       size_t((vl->tileCols + 2 * R) * (vl->tileCols + 2 * R) * (vl->tileK + 2 * R + 1)) * sizeof(int);
    for (int attempt = 0; attempt < 2; ++attempt)
    {
        vl->nbr16.ensure(ctx, int64_t(vl->capFull) * vl->strideS);
        hipLaunchKernelGGL(k_zeroI32v, dim3(1), dim3(64), 0, ctx->stream, dmax, int64_t(1));
        if (vl->ntiles > 0 && end > begin)
        {
            auto launch = [&](auto kernel) {
                hipLaunchKernelGGL(kernel, dim3(unsigned(vl->ntiles)), dim3(vl->block), lds, ctx->stream,
                                   vl->bin.sortedPos.p, vl->bin.sorted.p, vl->bin.cellOf.p, vl->bin.offsets.p,
                                   vl->tiles.p, geom(vl), g, begin, end, rn * rn, vl->nbr16.p, vl->strideS,
                                   vl->capFull, vl->cnt16.p, dmax, stagedCap);
            };
            const bool wide = int64_t(vl->capFull) * vl->strideS >= (int64_t(1) << 31);
            if (R == 1)
                wide ? launch(k_verletTiled<1, true>) : launch(k_verletTiled<1, false>);
            else
                wide ? launch(k_verletTiled<2, true>) : launch(k_verletTiled<2, false>);
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
    if (rows) *rows = vl->end - vl->begin;
    if (maxFull) *maxFull = vl->maxFull;
    if (totalFull)
    {
        auto* v = const_cast<mrmd_verlet_list*>(vl);
        const int64_t r = v->end - v->begin;
        std::vector<int32_t> h(static_cast<size_t>(r));
        if (r)
        {
            try
            {
                v->ensureGlobal();
            }
            catch (...)
            {
                return MRMD_ERR_HIP;
            }
            if (hipMemcpy(h.data(), v->cnt.p, size_t(r) * 4, hipMemcpyDeviceToHost) != hipSuccess) return MRMD_ERR_HIP;
        }
        int64_t t = 0;
        for (auto c : h) t += c;
        *totalFull = t;
    }
    if (dims)
        for (int d = 0; d < 3; ++d) dims[d] = vl->bin.grid.n[d];
    return MRMD_OK;
}

}  // extern "C"
