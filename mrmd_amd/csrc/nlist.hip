// Linked cell list, cell permutation and the Verlet list build on the device.
//
// Binning is a stable counting sort: per-record cell index, per-cell counts (atomics), a
// device scan, an atomic slot fill and a per-cell insertion sort that restores ascending
// record order inside each cell (Cabana's own fill order is nondeterministic).
//
// The Verlet build runs one thread per row over the 27-cell stencil with Cabana's
// min-distance prune and writes BOTH directions of every pair into a column-major ELL
// table (lanes of a wave write/read consecutive words).  Bit 31 of an entry marks the
// pairs Cabana's half list keeps (position-lexicographic discriminator), so the half list
// is an exact sub-list and force kernels can choose between the atomic-free full rows and
// the reference's half rows.
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
__global__ void k_binCount(const double4* pos, int64_t begin, int64_t n, GridD g, int32_t* cellOf, int32_t* counts)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double4 p = pos[begin + k];
    const int c = g.cell(p.x, p.y, p.z);
    cellOf[k] = c;
    atomicAdd(&counts[c], 1);
}

__global__ void k_binFill(const int32_t* cellOf, int64_t begin, int64_t n, const int32_t* offsets, int32_t* fill,
                          int32_t* sorted)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = cellOf[k];
    const int slot = offsets[c] + atomicAdd(&fill[c], 1);
    sorted[slot] = int32_t(begin + k);
}

// restore ascending record order inside each cell (insertion sort; cells hold tens of
// records, a heap sort covers pathological crowding)
__global__ void k_binSortCells(const int32_t* offsets, const int32_t* counts, int32_t* sorted, int ncell)
{
    const int c = int(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    int32_t* a = sorted + offsets[c];
    const int n = counts[c];
    if (n <= 96)
    {
        for (int i = 1; i < n; ++i)
        {
            const int32_t v = a[i];
            int j = i - 1;
            while (j >= 0 && a[j] > v)
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
            if (child + 1 < end && a[child] < a[child + 1]) ++child;
            if (a[root] < a[child])
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

void Binning::build(mrmd_ctx* ctx, const double4* pos, int64_t b, int64_t e, const GridD& g, bool withPos)
{
    grid = g;
    begin = b;
    end = e;
    const int64_t n = e - b;
    const int ncell = g.total();
    cellOf.reserve(ctx, std::max<int64_t>(growCap(n), 1));
    sorted.reserve(ctx, std::max<int64_t>(growCap(n), 1));
    counts.reserve(ctx, ncell + 1);
    offsets.reserve(ctx, ncell + 1);
    fill.reserve(ctx, ncell + 1);
    hipLaunchKernelGGL(k_zeroI32, dim3(gridFor(ncell)), dim3(BLOCK), 0, ctx->stream, counts.p, int64_t(ncell));
    hipLaunchKernelGGL(k_zeroI32, dim3(gridFor(ncell)), dim3(BLOCK), 0, ctx->stream, fill.p, int64_t(ncell));
    if (n > 0)
        hipLaunchKernelGGL(k_binCount, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, pos, b, n, g, cellOf.p,
                           counts.p);
    scanExclusiveI32(ctx, counts.p, offsets.p, ncell, nullptr);
    if (n > 0)
    {
        hipLaunchKernelGGL(k_binFill, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, cellOf.p, b, n, offsets.p,
                           fill.p, sorted.p);
        hipLaunchKernelGGL(k_binSortCells, dim3(gridFor(ncell)), dim3(BLOCK), 0, ctx->stream, offsets.p, counts.p,
                           sorted.p, ncell);
    }
    if (withPos && n > 0)
    {
        sortedPos.reserve(ctx, growCap(n));
        hipLaunchKernelGGL(k_gatherPos, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, pos, sorted.p, n,
                           sortedPos.p);
    }
    checkLaunch();
}

// =========================================================================================
namespace
{
// one thread per row; full rows with the half-list flag in bit 31
__global__ __launch_bounds__(BLOCK) void k_verletBuild(const double4* __restrict__ pos, GridD g, int range,
                                                       const int32_t* __restrict__ cellOf,
                                                       const int32_t* __restrict__ offsets,
                                                       const int32_t* __restrict__ counts,
                                                       const int32_t* __restrict__ sorted,
                                                       const double4* __restrict__ sortedPos, int64_t begin,
                                                       int64_t nrows, double rsq, uint32_t* nbr, int64_t stride,
                                                       int cap, int32_t* cnt, int32_t* maxCnt)
{
    const int64_t r = int64_t(xcdBlock()) * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int64_t p = begin + r;
    const double4 xp = pos[p];
    const int cell = cellOf[p];
    const int ci = cell / (g.n[1] * g.n[2]);
    const int cj = (cell / g.n[2]) % g.n[1];
    const int ck = cell % g.n[2];
    const int imin = max(ci - range, 0), imax = min(ci + range + 1, g.n[0]);
    const int jmin = max(cj - range, 0), jmax = min(cj + range + 1, g.n[1]);
    const int kmin = max(ck - range, 0), kmax = min(ck + range + 1, g.n[2]);
    int c = 0;
    for (int i = imin; i < imax; ++i)
        for (int j = jmin; j < jmax; ++j)
            for (int k = kmin; k < kmax; ++k)
            {
                if (g.minDistSqr(xp.x, xp.y, xp.z, i, j, k) > rsq) continue;
                const int cc = (i * g.n[1] + j) * g.n[2] + k;
                const int s0 = offsets[cc], s1 = s0 + counts[cc];
                for (int s = s0; s < s1; ++s)
                {
                    const int32_t n = sorted[s];
                    if (n == p) continue;
                    const double4 xn = sortedPos[s];
                    const double dx = xp.x - xn.x;
                    const double dy = xp.y - xn.y;
                    const double dz = xp.z - xn.z;
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 <= rsq)
                    {
                        const bool half = (xn.x > xp.x) ||
                                          ((xn.x == xp.x) && ((xn.y > xp.y) || ((xn.y == xp.y) && (xn.z > xp.z))));
                        if (c < cap) nbr[int64_t(c) * stride + r] = uint32_t(n) | (half ? 0x80000000u : 0u);
                        ++c;
                    }
                }
            }
    cnt[r] = c;
    atomicMax(maxCnt, c);
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

void verletBuild(mrmd_verlet_list* vl, const double4* pos, int64_t nall, int64_t begin, int64_t end, double rn,
                 double ratio, const double* gmin, const double* gmax, int64_t maxNeigh)
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
    const double gs = ratio * rn;
    const double delta[3] = {gs, gs, gs};
    const GridD g = makeGrid(gmin, gmax, delta);
    vl->bin.build(ctx, pos, 0, nall, g, true);
    const int64_t rows = end - begin;
    vl->stride = std::max<int64_t>(64, (rows + 63) / 64 * 64);
    vl->cnt.reserve(ctx, vl->stride);
    vl->halfValid = false;
    if (vl->capFull <= 0) vl->capFull = int32_t(std::max<int64_t>(2 * maxNeigh + 16, 96));
    int32_t* dmax = static_cast<int32_t*>(ctxScratch(ctx, 2, 64));
    for (int attempt = 0; attempt < 2; ++attempt)
    {
        vl->nbr.reserve(ctx, int64_t(vl->capFull) * vl->stride);
        hipLaunchKernelGGL(k_zeroI32, dim3(1), dim3(64), 0, ctx->stream, dmax, int64_t(1));
        if (rows > 0)
            hipLaunchKernelGGL(k_verletBuild, dim3(gridFor(rows)), dim3(BLOCK), 0, ctx->stream, pos, g, vl->range,
                               vl->bin.cellOf.p, vl->bin.offsets.p, vl->bin.counts.p, vl->bin.sorted.p,
                               vl->bin.sortedPos.p, begin, rows, rn * rn, vl->nbr.p, vl->stride, vl->capFull,
                               vl->cnt.p, dmax);
        checkLaunch();
        int32_t hmax = 0;
        HIPCHECK(hipMemcpyAsync(&hmax, dmax, 4, hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
        vl->maxFull = hmax;
        if (hmax <= vl->capFull) break;
        // Cabana VerletLayout2D: re-count and refill with the larger width
        vl->capFull = hmax + hmax / 8 + 8;
    }
    if (vl->maxFull > vl->capFull) fail(MRMD_ERR_STATE, "neighbour rows did not fit after a re-count");
}
}  // namespace

void mrmd_verlet_list::ensureHalf()
{
    if (halfValid) return;
    const int64_t rows = end - begin;
    capHalf = std::max(maxFull, 1);
    hnbr.reserve(ctx, int64_t(capHalf) * stride);
    hcnt.reserve(ctx, stride);
    if (rows > 0)
        hipLaunchKernelGGL(k_halfExtract, dim3(gridFor(rows)), dim3(BLOCK), 0, ctx->stream, nbr.p, cnt.p, rows, stride,
                           hnbr.p, hcnt.p);
    checkLaunch();
    halfValid = true;
}

// =========================================================================================
namespace
{
// Cabana::permute: new[begin + k] = old[perm[k]] for every field
__global__ void k_permuteGather(AtomsView src, AtomsView dst, const int32_t* perm, int64_t n, int64_t begin)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t s = perm[k], d = begin + k;
    dst.pos[d] = src.pos[s];
    for (int c = 0; c < 3; ++c)
    {
        dst.vel[c * dst.cap + d] = src.vel[c * src.cap + s];
        dst.force[c * dst.cap + d] = src.force[c * src.cap + s];
    }
    dst.mass[d] = src.mass[s];
    dst.charge[d] = src.charge[s];
    dst.relmass[d] = src.relmass[s];
}

template <class T>
__global__ void k_copyBack(T* dst, const T* src, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}
}  // namespace

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
    const int64_t n = cl->bin.end - cl->bin.begin;
    REQUIRE(cl->bin.end <= a->size, "cell list covers records beyond the atoms");
    if (n == 0) return MRMD_OK;
    mrmd_ctx* ctx = a->ctx;
    KTimer t(ctx, "atoms_permute");
    // gather into scratch buffers of the same capacity, then swap them in
    mrmd_atoms tmp;
    tmp.ctx = ctx;
    tmp.cap = a->cap;
    tmp.pos.reserve(ctx, a->cap);
    tmp.vel.reserve(ctx, 3 * a->cap);
    tmp.force.reserve(ctx, 3 * a->cap);
    tmp.mass.reserve(ctx, a->cap);
    tmp.charge.reserve(ctx, a->cap);
    tmp.relmass.reserve(ctx, a->cap);
    hipLaunchKernelGGL(k_permuteGather, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, a->view(), tmp.view(),
                       cl->bin.sorted.p, n, cl->bin.begin);
    // records outside [begin, end) are unchanged: copy them across
    const int64_t b = cl->bin.begin, e = cl->bin.end, s = a->size;
    auto copyRange = [&](int64_t lo, int64_t hi)
    {
        if (hi <= lo) return;
        HIPCHECK(hipMemcpyAsync(tmp.pos.p + lo, a->pos.p + lo, size_t(hi - lo) * 32, hipMemcpyDeviceToDevice, ctx->stream));
        for (int c = 0; c < 3; ++c)
        {
            HIPCHECK(hipMemcpyAsync(tmp.vel.p + c * a->cap + lo, a->vel.p + c * a->cap + lo, size_t(hi - lo) * 8,
                                    hipMemcpyDeviceToDevice, ctx->stream));
            HIPCHECK(hipMemcpyAsync(tmp.force.p + c * a->cap + lo, a->force.p + c * a->cap + lo, size_t(hi - lo) * 8,
                                    hipMemcpyDeviceToDevice, ctx->stream));
        }
        HIPCHECK(hipMemcpyAsync(tmp.mass.p + lo, a->mass.p + lo, size_t(hi - lo) * 8, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHECK(hipMemcpyAsync(tmp.charge.p + lo, a->charge.p + lo, size_t(hi - lo) * 8, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHECK(hipMemcpyAsync(tmp.relmass.p + lo, a->relmass.p + lo, size_t(hi - lo) * 8, hipMemcpyDeviceToDevice,
                                ctx->stream));
    };
    copyRange(0, b);
    copyRange(e, s);
    checkLaunch();
    sync(ctx);
    std::swap(a->pos.p, tmp.pos.p);
    std::swap(a->vel.p, tmp.vel.p);
    std::swap(a->force.p, tmp.force.p);
    std::swap(a->mass.p, tmp.mass.p);
    std::swap(a->charge.p, tmp.charge.p);
    std::swap(a->relmass.p, tmp.relmass.p);
    MRMD_API_END
}

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

int mrmd_verlet_build(mrmd_verlet_list* vl, const mrmd_atoms* a, int64_t begin, int64_t end, double rn, double ratio,
                      const double* gmin, const double* gmax, int64_t maxNeigh)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl && a && gmin && gmax, "invalid arguments");
    verletBuild(vl, a->pos.p, a->nlocal + a->nghost, begin, end, rn, ratio, gmin, gmax, maxNeigh);
    MRMD_API_END
}

int mrmd_verlet_build_molecules(mrmd_verlet_list* vl, const mrmd_molecules* m, int64_t begin, int64_t end, double rn,
                                double ratio, const double* gmin, const double* gmax, int64_t maxNeigh)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl && m && gmin && gmax, "invalid arguments");
    verletBuild(vl, m->pos.p, m->nlocal + m->nghost, begin, end, rn, ratio, gmin, gmax, maxNeigh);
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
    int32_t* dev = nullptr;
    HIPCHECK(hipMalloc(&dev, size_t(rows * rowStride) * 4));
    hipLaunchKernelGGL(k_rowsToHost, dim3(gridFor(rows)), dim3(BLOCK), 0, ctx->stream,
                       reinterpret_cast<const uint32_t*>(vl->hnbr.p), vl->hcnt.p, rows, vl->stride, dev, rowStride, 0);
    HIPCHECK(hipMemcpyAsync(host, dev, size_t(rows * rowStride) * 4, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    (void)hipFree(dev);
    MRMD_API_END
}

int mrmd_verlet_full_neighbors(mrmd_verlet_list* vl, int32_t* host, int32_t* counts, int64_t n, int64_t rowStride)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl && rowStride > 0, "invalid arguments");
    mrmd_ctx* ctx = vl->ctx;
    const int64_t rows = std::min<int64_t>(n, vl->end - vl->begin);
    if (rows <= 0) return MRMD_OK;
    if (counts) HIPCHECK(hipMemcpyAsync(counts, vl->cnt.p, size_t(rows) * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (host)
    {
        int32_t* dev = nullptr;
        HIPCHECK(hipMalloc(&dev, size_t(rows * rowStride) * 4));
        hipLaunchKernelGGL(k_rowsToHost, dim3(gridFor(rows)), dim3(BLOCK), 0, ctx->stream, vl->nbr.p, vl->cnt.p, rows,
                           vl->stride, dev, rowStride, 1);
        HIPCHECK(hipMemcpyAsync(host, dev, size_t(rows * rowStride) * 4, hipMemcpyDeviceToHost, ctx->stream));
        sync(ctx);
        (void)hipFree(dev);
    }
    sync(ctx);
    MRMD_API_END
}

int mrmd_verlet_count_within(mrmd_verlet_list* vl, const mrmd_atoms* a, double r, int64_t* half, int64_t* full)
{
    MRMD_API_BEGIN(vl ? vl->ctx : nullptr)
    REQUIRE(vl && a && vl->nrows <= a->size, "invalid arguments");
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
        if (r && hipMemcpy(h.data(), v->cnt.p, size_t(r) * 4, hipMemcpyDeviceToHost) != hipSuccess) return MRMD_ERR_HIP;
        int64_t t = 0;
        for (auto c : h) t += c;
        *totalFull = t;
    }
    if (dims)
        for (int d = 0; d < 3; ++d) dims[d] = vl->bin.grid.n[d];
    return MRMD_OK;
}

}  // extern "C"
