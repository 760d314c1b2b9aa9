// Cell binning, the linked cell list and the cell permutation of the atoms.
//
// Binning is a stable counting sort: per-record cell index, per-cell counts (atomics), a
// device scan, an atomic slot fill and a per-cell insertion sort that restores ascending
// record order inside each cell (Cabana's own fill order is nondeterministic).
#include <algorithm>

#include "nlist.hpp"

using namespace mrmd_hip;

namespace mrmd_hip
{
GridD makeGrid(const double* gmin, const double* gmax, const double* delta)
{
    GridD g{};
    for (int d = 0; d < 3; ++d)
    {
        REQUIRE(gmax[d] > gmin[d] && delta[d] > 0.0, "invalid grid along %d", d);
        int n = int(std::floor((gmax[d] - gmin[d]) * (1.0 / delta[d])));
        if (n < 1) n = 1;
        g.n[d] = n;
        g.min[d] = gmin[d];
        g.dx[d] = (gmax[d] - gmin[d]) / n;
        g.rdx[d] = 1.0 / g.dx[d];
    }
    REQUIRE(double(g.n[0]) * g.n[1] * g.n[2] < 2e9, "cell grid too large");
    return g;
}
}  // namespace mrmd_hip

namespace
{
// WRAP: each position is first mapped into the box (PeriodicMapping) and written back; its
// current value is wb.from[wb.slotOf[i]] when wb.from is set
template <bool WRAP>
__global__ void k_binCount(const double4* pos, int64_t begin, int64_t n, GridD g, int32_t* cellOf, int32_t* counts,
                           double4* wrapPos, WrapBox wb)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = k < n;
    int c = -1;
    if (valid)
    {
        double4 p = (WRAP && wb.from) ? wb.from[wb.slotOf[begin + k]] : pos[begin + k];
        if (WRAP)
        {
            wrapInto(p, wb.min, wb.max, wb.diam);
            wrapPos[begin + k] = p;
        }
        c = g.cell(p.x, p.y, p.z);
        cellOf[k] = c;
    }
    const Run r = laneRun(c, valid);
    if (valid && r.head == int(threadIdx.x) % WAVE) atomicAdd(&counts[c], r.len);
}

__global__ void k_binFill(const int32_t* cellOf, int64_t begin, int64_t n, const int32_t* offsets, int32_t* fill,
                          int32_t* sorted)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = k < n;
    const int c = valid ? cellOf[k] : -1;
    const Run r = laneRun(c, valid);
    const int lane = int(threadIdx.x) % WAVE;
    int base = 0;
    if (valid && r.head == lane) base = atomicAdd(&fill[c], r.len);
    base = __shfl(base, r.head, WAVE);
    if (valid) sorted[offsets[c] + base + (lane - r.head)] = int32_t(begin + k);
}

// order the records inside each cell by ascending record index.  A workgroup takes 128
// consecutive cells, whose records are one contiguous run of the sorted array, and loads the
// run into LDS once; every record's place in its cell is its rank there, the count of the
// cell's records with a smaller index (one lane per record; the lanes of one cell read the
// same record at a time), so the records go straight to their places.  A run longer than
// the LDS buffer (pathological crowding) sorts in place in global memory, one lane per cell,
// with a heap sort for crowded cells.  (The z order of the list build is k_sortCellsZ's.)
constexpr int SORT_RUN = 4096;
constexpr int SORT_CELLS = 128;  // cells per workgroup: ~20 records each fill a run of ~2500

__global__ __launch_bounds__(SORT_CELLS) void k_binSortCells(const int32_t* offsets, const int32_t* counts, int32_t* sorted,
                                                        int ncell)
{
    __shared__ int32_t idx[SORT_RUN];
    __shared__ uint8_t cellOfRec[SORT_RUN];
    __shared__ int32_t cellStart[SORT_CELLS], cellCount[SORT_CELLS];
    const int c0 = int(blockIdx.x) * SORT_CELLS;
    const int c1 = min(c0 + SORT_CELLS, ncell);
    const int r0 = offsets[c0], r1 = offsets[c1 - 1] + counts[c1 - 1];
    const int c = c0 + int(threadIdx.x);
    if (r1 - r0 <= SORT_RUN)
    {
        // eight independent loads per lane in flight
        const int len = r1 - r0;
        for (int k0 = threadIdx.x; k0 < len; k0 += 8 * SORT_CELLS)
        {
            int32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = k0 + u * SORT_CELLS < len ? sorted[r0 + k0 + u * SORT_CELLS] : 0;
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (k0 + u * SORT_CELLS < len) idx[k0 + u * SORT_CELLS] = v[u];
        }
        if (c < c1)
        {
            const int o = offsets[c] - r0, n = counts[c];
            cellStart[threadIdx.x] = o;
            cellCount[threadIdx.x] = n;
            for (int i = 0; i < n; ++i) cellOfRec[o + i] = uint8_t(threadIdx.x);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < r1 - r0; k += SORT_CELLS)
        {
            const int32_t v = idx[k];
            const int cl = cellOfRec[k];
            const int o = cellStart[cl], n = cellCount[cl];
            int rank = 0;
            for (int j = o; j < o + n; ++j) rank += idx[j] < v ? 1 : 0;
            sorted[r0 + o + rank] = v;
        }
        return;
    }
    if (c >= c1) return;
    int32_t* a = sorted + offsets[c];
    const int n = counts[c];
    auto lessRec = [](int32_t x, int32_t y) { return x < y; };
    if (n <= 96)
    {
        for (int i = 1; i < n; ++i)
        {
            const int32_t v = a[i];
            int j = i - 1;
            while (j >= 0 && lessRec(v, a[j]))
            {
                a[j + 1] = a[j];
                --j;
            }
            a[j + 1] = v;
        }
        return;
    }
    // heap sort
    auto sift = [&](int start, int end)
    {
        int root = start;
        while (2 * root + 1 < end)
        {
            int child = 2 * root + 1;
            if (child + 1 < end && lessRec(a[child], a[child + 1])) ++child;
            if (lessRec(a[root], a[child]))
            {
                const int32_t t = a[root];
                a[root] = a[child];
                a[child] = t;
                root = child;
            }
            else
                return;
        }
    };
    for (int s = n / 2 - 1; s >= 0; --s) sift(s, n);
    for (int end = n - 1; end > 0; --end)
    {
        const int32_t t = a[0];
        a[0] = a[end];
        a[end] = t;
        sift(0, end);
    }
}

// (z, index) order inside each cell: one group of 32 lanes per cell (sortCellByZ)
// outPos (if not null): each slot's record copied there once the cell is in order (the
// sorted-slot positions), and slotOf (if not null) the slot of each record
__global__ __launch_bounds__(BLOCK) void k_sortCellsZ(const int32_t* offsets, const int32_t* counts, int32_t* sorted,
                                                      int ncell, const double4* __restrict__ pos, double4* outPos,
                                                      int32_t* slotOf)
{
    const int64_t c = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) / SORT_GROUP;
    if (c >= ncell) return;
    const int lane = int(threadIdx.x % SORT_GROUP);
    const int base = offsets[c], n = counts[c];
    int32_t* recs = sorted + base;
    sortCellByZ(recs, n, pos, lane);
    if (!outPos) return;
    // every lane reads back only the places it wrote itself; a crowded cell, sorted by lane 0
    // alone in memory, is copied by lane 0 alone
    const bool serial = n > 4 * SORT_GROUP;
    if (serial && lane != 0) return;
    for (int e = serial ? 0 : lane; e < n; e += serial ? 1 : SORT_GROUP)
    {
        const int32_t v = recs[e];
        outPos[base + e] = pos[v];
        if (slotOf) slotOf[v] = base + e;
    }
}

__global__ void k_gatherPos(const double4* pos, const int32_t* sorted, int64_t n, double4* out)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) out[k] = pos[sorted[k]];
}

__global__ void k_zeroI32(int32_t* p, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}
}  // namespace

void Binning::build(mrmd_ctx* ctx, const double4* pos, int64_t b, int64_t e, const GridD& g, bool withPos,
                    bool zOrder, double4* wrapPos, const WrapBox* wrap, int32_t* slotOfOut)
{
    grid = g;
    begin = b;
    end = e;
    const int64_t n = e - b;
    const int ncell = g.total();
    cellOf.ensure(ctx, std::max<int64_t>(n, 1));
    sorted.ensure(ctx, std::max<int64_t>(n, 1));
    counts.ensure(ctx, ncell + 1);
    offsets.ensure(ctx, ncell + 1);
    fill.ensure(ctx, ncell + 1);
    hipLaunchKernelGGL(k_zeroI32, dim3(gridFor(ncell)), dim3(BLOCK), 0, ctx->stream, counts.p, int64_t(ncell));
    hipLaunchKernelGGL(k_zeroI32, dim3(gridFor(ncell)), dim3(BLOCK), 0, ctx->stream, fill.p, int64_t(ncell));
    if (n > 0)
    {
        if (wrap)
            hipLaunchKernelGGL(k_binCount<true>, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, pos, b, n, g, cellOf.p,
                               counts.p, wrapPos, *wrap);
        else
            hipLaunchKernelGGL(k_binCount<false>, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, pos, b, n, g,
                               cellOf.p, counts.p, nullptr, WrapBox{});
    }
    scanExclusiveI32(ctx, counts.p, offsets.p, ncell, nullptr);
    if (n > 0)
    {
        hipLaunchKernelGGL(k_binFill, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, cellOf.p, b, n, offsets.p,
                           fill.p, sorted.p);
        if (withPos) sortedPos.ensure(ctx, n);
        if (zOrder)
            // the z sort also writes the sorted-slot positions (and the slot of each record)
            hipLaunchKernelGGL(k_sortCellsZ, dim3(unsigned((int64_t(ncell) * SORT_GROUP + BLOCK - 1) / BLOCK)),
                               dim3(BLOCK), 0, ctx->stream, offsets.p, counts.p, sorted.p, ncell, pos,
                               withPos ? sortedPos.p : nullptr, withPos ? slotOfOut : nullptr);
        else
            hipLaunchKernelGGL(k_binSortCells, dim3((ncell + SORT_CELLS - 1) / SORT_CELLS), dim3(SORT_CELLS), 0,
                               ctx->stream, offsets.p, counts.p, sorted.p, ncell);
        if (withPos && !zOrder)
            hipLaunchKernelGGL(k_gatherPos, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, pos, sorted.p, n,
                               sortedPos.p);
    }
    checkLaunch();
}

// =========================================================================================
namespace
{
// every field of records [lo, hi) copied unchanged (the records a permutation leaves alone)
// scalars: bit 0 mass, bit 1 charge, bit 2 relmass are moved (a field of one value over all
// records stays where it is)
__global__ void k_copyRecords(AtomsView src, AtomsView dst, int64_t lo, int64_t hi, unsigned scalars)
{
    const int64_t i = lo + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= hi) return;
    dst.pos[i] = src.pos[i];
    for (int c = 0; c < 3; ++c)
    {
        dst.vel[c * dst.cap + i] = src.vel[c * src.cap + i];
        dst.force[c * dst.cap + i] = src.force[c * src.cap + i];
    }
    if (scalars & 1u) dst.mass[i] = src.mass[i];
    if (scalars & 2u) dst.charge[i] = src.charge[i];
    if (scalars & 4u) dst.relmass[i] = src.relmass[i];
}

// Cabana::permute: new[begin + k] = old[perm[k]] for every field (FORCE: forces too)
// (scalars as for k_copyRecords)
template <bool FORCE>
__global__ void k_permuteGather(AtomsView src, AtomsView dst, const int32_t* perm, int64_t n, int64_t begin,
                                unsigned scalars)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t s = perm[k], d = begin + k;
    dst.pos[d] = src.pos[s];
    for (int c = 0; c < 3; ++c)
    {
        dst.vel[c * dst.cap + d] = src.vel[c * src.cap + s];
        if (FORCE) dst.force[c * dst.cap + d] = src.force[c * src.cap + s];
    }
    if (scalars & 1u) dst.mass[d] = src.mass[s];
    if (scalars & 2u) dst.charge[d] = src.charge[s];
    if (scalars & 4u) dst.relmass[d] = src.relmass[s];
}

template <class T>
__global__ void k_copyBack(T* dst, const T* src, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}
}  // namespace

namespace mrmd_hip
{
// PeriodicMapping of the real atoms fused with the cell list build over them (the LJ time
// loop's rebuild: the wrap and the binning's pass over the positions are one pass)
void cellListBuildWrapped(mrmd_cell_list* cl, mrmd_atoms* a, const mrmd_box* box, const double* delta,
                          const double4* from, const int32_t* slotOf)
{
    KTimer t(cl->ctx, "cell_list_build");
    WrapBox wb;
    for (int d = 0; d < 3; ++d)
    {
        wb.min[d] = box->min[d];
        wb.max[d] = box->max[d];
        wb.diam[d] = box->max[d] - box->min[d];
    }
    wb.from = from;
    wb.slotOf = slotOf;
    cl->bin.build(cl->ctx, a->pos.p, 0, a->nlocal, makeGrid(box->min, box->max, delta), false, false, a->pos.p, &wb);
}

// Atoms::permute by a cell list's order; full = false (the time loop, right before the ghosts
// are made again and the forces computed again) moves neither the forces of the permuted
// records nor the records outside the list's range, whose values are rewritten before use
void permuteAtoms(mrmd_atoms* a, const mrmd_cell_list* cl, bool full)
{
    const int64_t n = cl->bin.end - cl->bin.begin;
    REQUIRE(cl->bin.end <= a->size, "cell list covers records beyond the atoms");
    if (n == 0) return;
    mrmd_ctx* ctx = a->ctx;
    KTimer t(ctx, "atoms_permute");
    // gather into the spare buffers (same capacity), then swap them in; the old buffers
    // become the spares of the next permutation.  A mass, charge or relmass field of one
    // value over all records is the same after any permutation: it is neither moved nor
    // swapped
    const unsigned scalars = (a->massUniform ? 0u : 1u) | (a->chargeUniform ? 0u : 2u) | (a->relmassUniform ? 0u : 4u);
    a->sparePos.reserve(ctx, a->cap);
    a->spareVel.reserve(ctx, 3 * a->cap);
    a->spareForce.reserve(ctx, 3 * a->cap);
    if (scalars & 1u) a->spareMass.reserve(ctx, a->cap);
    if (scalars & 2u) a->spareCharge.reserve(ctx, a->cap);
    if (scalars & 4u) a->spareRelmass.reserve(ctx, a->cap);
    if (full)
        hipLaunchKernelGGL(k_permuteGather<true>, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, a->view(),
                           a->spareView(), cl->bin.sorted.p, n, cl->bin.begin, scalars);
    else
        hipLaunchKernelGGL(k_permuteGather<false>, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, a->view(),
                           a->spareView(), cl->bin.sorted.p, n, cl->bin.begin, scalars);
    // records outside [begin, end) are unchanged: copy them across
    const int64_t b = cl->bin.begin, e = cl->bin.end, s = a->size;
    auto copyRange = [&](int64_t lo, int64_t hi)
    {
        if (hi <= lo) return;
        hipLaunchKernelGGL(k_copyRecords, dim3(gridFor(hi - lo)), dim3(BLOCK), 0, ctx->stream, a->view(),
                           a->spareView(), lo, hi, scalars);
    };
    if (full)
    {
        copyRange(0, b);
        copyRange(e, s);
    }
    checkLaunch();
    a->pos.swap(a->sparePos);
    a->vel.swap(a->spareVel);
    a->force.swap(a->spareForce);
    if (scalars & 1u) a->mass.swap(a->spareMass);
    if (scalars & 2u) a->charge.swap(a->spareCharge);
    if (scalars & 4u) a->relmass.swap(a->spareRelmass);
}
}  // namespace mrmd_hip

extern "C" {

int mrmd_cell_list_create(mrmd_ctx* c, mrmd_cell_list** out)
{
    MRMD_API_BEGIN(c)
    REQUIRE(out, "null output");
    auto* cl = new mrmd_cell_list();
    cl->ctx = c;
    *out = cl;
    MRMD_API_END
}

int mrmd_cell_list_destroy(mrmd_cell_list* cl)
{
    if (!cl) return MRMD_OK;
    (void)hipStreamSynchronize(cl->ctx->stream);
    delete cl;
    return MRMD_OK;
}

int mrmd_cell_list_build(mrmd_cell_list* cl, const mrmd_atoms* a, int64_t begin, int64_t end, const double* delta,
                         const double* gmin, const double* gmax)
{
    MRMD_API_BEGIN(cl ? cl->ctx : nullptr)
    REQUIRE(cl && a && delta && gmin && gmax, "invalid arguments");
    REQUIRE(begin >= 0 && begin <= end && end <= a->size, "invalid range");
    KTimer t(cl->ctx, "cell_list_build");
    cl->bin.build(cl->ctx, a->pos.p, begin, end, makeGrid(gmin, gmax, delta), false);
    MRMD_API_END
}

int mrmd_cell_list_dims(const mrmd_cell_list* cl, int32_t* dims)
{
    if (!cl || !dims) return MRMD_ERR_ARG;
    for (int d = 0; d < 3; ++d) dims[d] = cl->bin.grid.n[d];
    return MRMD_OK;
}

int mrmd_cell_list_download(const mrmd_cell_list* clc, int32_t* cellOf, int32_t* counts, int64_t* offsets,
                            int64_t* perm)
{
    auto* cl = const_cast<mrmd_cell_list*>(clc);
    MRMD_API_BEGIN(cl ? cl->ctx : nullptr)
    REQUIRE(cl, "null cell list");
    mrmd_ctx* ctx = cl->ctx;
    const int64_t n = cl->bin.end - cl->bin.begin;
    const int ncell = cl->bin.grid.total();
    if (cellOf && n) HIPCHECK(hipMemcpyAsync(cellOf, cl->bin.cellOf.p, size_t(n) * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (counts) HIPCHECK(hipMemcpyAsync(counts, cl->bin.counts.p, size_t(ncell) * 4, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<int32_t> off(static_cast<size_t>(ncell)), pm(static_cast<size_t>(n));
    if (offsets) HIPCHECK(hipMemcpyAsync(off.data(), cl->bin.offsets.p, size_t(ncell) * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (perm && n) HIPCHECK(hipMemcpyAsync(pm.data(), cl->bin.sorted.p, size_t(n) * 4, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    if (offsets)
        for (int i = 0; i < ncell; ++i) offsets[i] = off[size_t(i)];
    if (perm)
        for (int64_t i = 0; i < n; ++i) perm[i] = pm[size_t(i)];
    MRMD_API_END
}

int mrmd_atoms_permute(mrmd_atoms* a, const mrmd_cell_list* cl)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && cl, "invalid arguments");
    permuteAtoms(a, cl, true);
    MRMD_API_END
}

}  // extern "C"
