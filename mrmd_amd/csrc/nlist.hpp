// Cell grid and neighbour-list device structures.
#pragma once

#include "containers.hpp"

namespace mrmd_hip
{
// Cartesian cell grid with Cabana LinkedCellList semantics (third-party Cabana 0.7
// Impl::CartesianGrid): n = floor((max-min) * (1/delta)), dx = (max-min)/n, cell index
// floor((x - min) * (1/dx)) with the outer boundary folded into the last cell and
// cardinal index (i * ny + j) * nz + k.  All arithmetic is unfused.
struct GridD
{
    double min[3], dx[3], rdx[3];
    int n[3];

    __host__ __device__ int locate1(double x, int d) const
    {
        int c = int(floor((x - min[d]) * rdx[d]));
        if (c == n[d]) c = n[d] - 1;
        // positions outside the grid are undefined in the reference; clamp instead
        c = c < 0 ? 0 : c;
        c = c >= n[d] ? n[d] - 1 : c;
        return c;
    }
    __host__ __device__ int cell(double x, double y, double z) const
    {
        return (locate1(x, 0) * n[1] + locate1(y, 1)) * n[2] + locate1(z, 2);
    }
    __host__ __device__ int total() const { return n[0] * n[1] * n[2]; }
    // Cabana minDistanceToPoint: squared distance from p to the cell box (0 inside)
    __host__ __device__ double minDistSqr(double px, double py, double pz, int i, int j, int k) const
    {
        const double p[3] = {px, py, pz};
        const int ijk[3] = {i, j, k};
        double r2 = 0.0;
        for (int d = 0; d < 3; ++d)
        {
            const double c = min[d] + (double(ijk[d]) + 0.5) * dx[d];
            double r = fabs(p[d] - c) - 0.5 * dx[d];
            r = r > 0.0 ? r : 0.0;
            r2 = r2 + r * r;
        }
        return r2;
    }
};

GridD makeGrid(const double* gmin, const double* gmax, const double* delta);

// binning of records [begin, end) of a double4 position array into a grid; the sorted
// order is ascending record index inside every cell (a stable counting sort), or with
// zOrder ascending z (record index breaking ties), which makes every column of cells one
// z-sorted run
struct Binning
{
    GridD grid{};
    int64_t begin = 0, end = 0;
    DBuf<int32_t> cellOf;   ///< cell of record begin+k
    DBuf<int32_t> counts;   ///< records per cell
    DBuf<int32_t> offsets;  ///< exclusive scan of counts
    DBuf<int32_t> fill;     ///< scratch cursor per cell
    DBuf<int32_t> sorted;   ///< slot -> record index
    DBuf<double4> sortedPos;

    /// wrap (if given): every position is first mapped into that box and written to
    /// wrapPos (the same array as pos for an in-place wrap); slotOfOut (z order with
    /// positions, begin 0): the slot of each record, written by the z sort
    void build(mrmd_ctx* ctx, const double4* pos, int64_t b, int64_t e, const GridD& g, bool withPos,
               bool zOrder = false, double4* wrapPos = nullptr, const WrapBox* wrap = nullptr,
               int32_t* slotOfOut = nullptr);
};
}  // namespace mrmd_hip

struct mrmd_cell_list
{
    mrmd_hip::CtxPtr ctx;
    mrmd_hip::Binning bin;
};

namespace mrmd_hip
{
// Atoms::permute by the cell list's order (full = false: the time loop's form, which leaves
// the permuted records' forces and the records outside the list's range unset, as the loop
// recomputes both before reading them)
void permuteAtoms(mrmd_atoms* a, const mrmd_cell_list* cl, bool full);
// PeriodicMapping of the real atoms and the cell list build over them in one pass
void cellListBuildWrapped(mrmd_cell_list* cl, mrmd_atoms* a, const mrmd_box* box, const double* delta,
                          const double4* from = nullptr, const int32_t* slotOf = nullptr);
}

namespace mrmd_hip
{
constexpr int MAX_RANGES = 25;  // (a + 2R)^2 staged columns: a <= 3 with R = 1, a = 1 with R = 2
constexpr int MAX_SELF = 9;     // a x a columns of rows
constexpr int STAGE_MAX = 2048; // staged records per tile (double4 -> 64 KB of LDS)

// One tile = a run of K cells along z in a x a neighbouring (i, j) columns.  Its rows are
// the sorted slots of those cells (one contiguous run per column); its neighbourhood is
// the (a + 2R)^2 columns around them, each a contiguous range of sorted slots (empty
// outside the grid), staged back to back in LDS at base[r].
struct TileInfo
{
    int32_t nr, staged, nself, nrows;
    int32_t begin[MAX_RANGES];
    int32_t base[MAX_RANGES];
    int32_t selfRange[MAX_SELF], selfBegin[MAX_SELF], selfPrefix[MAX_SELF + 1];
};

// sorted slot of the t-th row of a tile (its columns' rows one column after the other)
__device__ inline int tileRow(const TileInfo& ti, int t, int& self)
{
    int r = 0;
    while (r + 1 < ti.nself && ti.selfPrefix[r + 1] <= t) ++r;
    self = r;
    return ti.selfBegin[r] + (t - ti.selfPrefix[r]);
}

// LDS slot of sorted slot s that lies in staged range r of the tile
__device__ inline int ldsSlot(const TileInfo& ti, int r, int s) { return ti.base[r] + (s - ti.begin[r]); }

// Neighbour entries are 16-bit LDS slots stored in chunks of 8 per row: entry k of sorted
// slot s sits at ((k / 8) * strideS + s) * 8 + k % 8, so one 16-byte load per lane brings
// 8 entries and a wave's load is 1 KB contiguous.
constexpr int CHUNK = 8;
__host__ __device__ inline int64_t entryIndex(int64_t k, int64_t s, int64_t strideS)
{
    return ((k >> 3) * strideS + s) * CHUNK + (k & 7);
}

// A staged LDS record: {x, y, z} (24 bytes) where the kernel needs positions only, the full
// {x, y, z, type} record (32 bytes) where it needs the type.  The smaller record fits more
// workgroups into a CU's LDS.
template <bool XYZ>
struct Staged;
template <>
struct Staged<true>
{
    static constexpr int bytes = 24;
    __device__ static void put(void* st, int q, const double4& v)
    {
        double* p = static_cast<double*>(st) + 3 * q;
        p[0] = v.x;
        p[1] = v.y;
        p[2] = v.z;
    }
    __device__ static double4 get(const void* st, int q)
    {
        const double* p = static_cast<const double*>(st) + 3 * q;
        return make_double4(p[0], p[1], p[2], 0.0);
    }
};
template <>
struct Staged<false>
{
    static constexpr int bytes = 32;
    __device__ static void put(void* st, int q, const double4& v) { static_cast<double4*>(st)[q] = v; }
    __device__ static double4 get(const void* st, int q) { return static_cast<const double4*>(st)[q]; }
};

// Rows are padded to whole chunks with a sentinel entry: the LDS slot just past the largest
// staged neighbourhood of the list, which every tiled force kernel fills with a record at
// (c, c, c), c = the largest upper grid bound + SENTINEL_GAP + 4 r_n (the list's
// `sentinel`): every point of the grid lies more than SENTINEL_GAP + 4 r_n from it along each
// axis, so a padding pair lies beyond any cutoff and needs no test of its own.
constexpr double SENTINEL_GAP = 1.0e3;
__host__ __device__ inline int sentinelSlot(int32_t maxStaged) { return maxStaged > 1 ? maxStaged : 1; }
template <bool XYZ>
__device__ inline void putSentinel(void* st, int32_t maxStaged, double coord, double w)
{
    if (threadIdx.x == 0) Staged<XYZ>::put(st, sentinelSlot(maxStaged), make_double4(coord, coord, coord, w));
}

// Sort the records of one cell by (z, record index) in a group of 32 lanes (half a wave):
// lane l holds records l, l + 32, ... (E per lane, E = 1 for cells of up to 32 records, 4 for
// up to 128; the places past the count hold +infinity), a bitonic network orders them (pairs
// 32 or more places apart inside a lane, nearer pairs across lanes by shuffles), and they are
// written back in place.  A cell of more than 128 records (pathological crowding) is
// insertion-sorted by lane 0 of the group alone, in global memory.  `n` is the same for the
// 32 lanes of a group.
constexpr int SORT_GROUP = 32;
// KEY: ties in z are broken by key[record] instead of the record index (records that move
// keep their key)
template <int E, bool KEY = false>
__device__ inline void sortGroup(int32_t* recs, int n, const double4* __restrict__ pos, int lane,
                                 const int32_t* __restrict__ key = nullptr)
{
    double z[E];
    int32_t v[E], k[E];
#pragma unroll
    for (int r = 0; r < E; ++r)
    {
        const int e = r * SORT_GROUP + lane;
        v[r] = e < n ? recs[e] : INT32_MAX;
        z[r] = e < n ? pos[v[r]].z : HUGE_VAL;
        if constexpr (KEY) k[r] = e < n ? key[v[r]] : INT32_MAX;
    }
    auto less = [](double za, int32_t va, double zb, int32_t vb) { return za < zb || (za == zb && va < vb); };
    auto tie = [&](int r) -> int32_t {
        if constexpr (KEY) return k[r];
        else return v[r];
    };
#pragma unroll
    for (int kk = 2; kk <= E * SORT_GROUP; kk <<= 1)
#pragma unroll
        for (int j = kk >> 1; j > 0; j >>= 1)
        {
            if (j >= SORT_GROUP)
            {
                // partners in one lane: register r against r ^ (j / 32)
                const int rj = j / SORT_GROUP;
#pragma unroll
                for (int r = 0; r < E; ++r)
                {
                    const int rp = r ^ rj;
                    if (rp < r) continue;
                    const int e = r * SORT_GROUP + lane;  // the lower place of the pair
                    const bool asc = (e & kk) == 0;
                    if (asc == less(z[rp], tie(rp), z[r], tie(r)))
                    {
                        const double tz = z[r];
                        const int32_t tv = v[r];
                        z[r] = z[rp];
                        v[r] = v[rp];
                        z[rp] = tz;
                        v[rp] = tv;
                        if constexpr (KEY)
                        {
                            const int32_t tk = k[r];
                            k[r] = k[rp];
                            k[rp] = tk;
                        }
                    }
                }
            }
            else
            {
#pragma unroll
                for (int r = 0; r < E; ++r)
                {
                    const int e = r * SORT_GROUP + lane;
                    const double zo = __shfl_xor(z[r], j, SORT_GROUP);
                    const int32_t vo = __shfl_xor(v[r], j, SORT_GROUP);
                    int32_t ko = vo;
                    if constexpr (KEY) ko = __shfl_xor(k[r], j, SORT_GROUP);
                    // the lower place of a pair keeps the smaller record in an ascending block
                    // (e & kk == 0) and the larger in a descending one, the upper place the other
                    const bool wantMin = ((e & kk) == 0) != ((e & j) != 0);
                    if (wantMin ? less(zo, ko, z[r], tie(r)) : less(z[r], tie(r), zo, ko))
                    {
                        z[r] = zo;
                        v[r] = vo;
                        if constexpr (KEY) k[r] = ko;
                    }
                }
            }
        }
#pragma unroll
    for (int r = 0; r < E; ++r)
    {
        const int e = r * SORT_GROUP + lane;
        if (e < n) recs[e] = v[r];
    }
}

template <bool KEY = false>
__device__ inline void sortCellByZ(int32_t* recs, int n, const double4* __restrict__ pos, int lane,
                                   const int32_t* __restrict__ key = nullptr)
{
    if (n <= 1) return;
    if (n <= SORT_GROUP)
    {
        sortGroup<1, KEY>(recs, n, pos, lane, key);
        return;
    }
    if (n <= 4 * SORT_GROUP)
    {
        sortGroup<4, KEY>(recs, n, pos, lane, key);
        return;
    }
    if (lane != 0) return;
    auto tieOf = [&](int32_t v) -> int32_t {
        if constexpr (KEY) return key[v];
        else return v;
    };
    for (int i = 1; i < n; ++i)
    {
        const int32_t v = recs[i];
        const double zv = pos[v].z;
        const int32_t tv = tieOf(v);
        int j = i - 1;
        while (j >= 0 && (zv < pos[recs[j]].z || (zv == pos[recs[j]].z && tv < tieOf(recs[j]))))
        {
            recs[j + 1] = recs[j];
            --j;
        }
        recs[j + 1] = v;
    }
}

// copy the tile's neighbourhood (contiguous runs of sorted slots) into LDS; the copy is
// unrolled so that four independent loads per lane are in flight
template <bool XYZ>
__device__ inline void stageTile(const TileInfo& ti, const double4* __restrict__ sortedPos, void* st)
{
    int r = 0;
    const int step = blockDim.x;
    auto source = [&](int q) -> int {
        if (q >= ti.staged) return -1;
        while (r + 1 < ti.nr && ti.base[r + 1] <= q) ++r;
        return ti.begin[r] + (q - ti.base[r]);
    };
    for (int q0 = threadIdx.x; q0 < ti.staged; q0 += 4 * step)
    {
        const int i0 = source(q0), i1 = source(q0 + step), i2 = source(q0 + 2 * step), i3 = source(q0 + 3 * step);
        double4 v0, v1, v2, v3;
        v0 = sortedPos[i0];
        if (i1 >= 0) v1 = sortedPos[i1];
        if (i2 >= 0) v2 = sortedPos[i2];
        if (i3 >= 0) v3 = sortedPos[i3];
        Staged<XYZ>::put(st, q0, v0);
        if (i1 >= 0) Staged<XYZ>::put(st, q0 + step, v1);
        if (i2 >= 0) Staged<XYZ>::put(st, q0 + 2 * step, v2);
        if (i3 >= 0) Staged<XYZ>::put(st, q0 + 3 * step, v3);
    }
}
}  // namespace mrmd_hip

struct mrmd_verlet_list
{
    mrmd_hip::CtxPtr ctx;
    int64_t nrows = 0;          ///< records of the binned slice (counts exist for all)
    int64_t begin = 0, end = 0; ///< rows = records [begin, end)
    double rn = 0, ratio = 1;
    int range = 1;
    int32_t capFull = 0;
    int32_t maxFull = 0;
    double sentinel = 0;        ///< coordinate of the padding entries' record (see putSentinel)
    bool forceWide = false;     ///< build with 64-bit entry offsets even for small lists
    int stageLimit = mrmd_hip::STAGE_MAX;  ///< staged records a tile may hold in LDS
    /// tiles whose neighbourhood exceeds stageLimit even at one cell column by one cell
    /// (dense inputs) stage it in a slice of global memory per tile instead: gstage holds
    /// ntiles x gstageRecs records of 32 bytes, which the list build and force kernels use
    /// in place of LDS (same entries, same bits)
    bool globalStage = false;
    int64_t gstageRecs = 0;
    mrmd_hip::DBuf<double4> gstage;

    // ---- tiled form (built on the device; what the force kernels read) ----
    int tileK = 0;              ///< cells per tile along z (0 = choose from the density)
    int tileCols = 2;           ///< a: tiles span a x a cell columns
    int block = 256;            ///< workgroup size of the tiled kernels
    int tilesPerCol = 0;
    int64_t ntiles = 0;
    int rowOrigin[3] = {0, 0, 0};  ///< first cell, per axis, of the box of cells holding rows
    int rowExtent[3] = {0, 0, 0};  ///< its extent in cells (tiles cover only this box)
    int32_t maxStaged = 0;
    int64_t strideS = 0;        ///< ELL stride over sorted slots (>= nrows, multiple of 64)
    mrmd_hip::DBuf<mrmd_hip::TileInfo> tiles;
    /// entries per sorted slot s: nbr16[entryIndex(k, s, strideS)] = LDS slot (15 bits) |
    /// half flag (bit 15)
    mrmd_hip::DBuf<uint16_t> nbr16;
    mrmd_hip::DBuf<int32_t> cnt16;  ///< count per sorted slot (0 for non-row records)
    /// The AdResS time loop builds only the rows that can interact before the next rebuild:
    /// with bandLo <= bandHi, a tile whose neighbourhood lies wholly outside the x band
    /// [bandLo, bandHi] (records that may carry a slab weight before the next rebuild) gets
    /// empty rows and tileSkip[t] = 1, which the tiled LJ_IdealGas kernel skips.  Such a list
    /// (`banded`) serves that kernel only; every other use asks for a full build.
    double bandLo = 1.0, bandHi = -1.0;
    bool banded = false;
    mrmd_hip::DBuf<int8_t> tileSkip;

    // ---- global-index form (rows = records begin+r), converted on demand ----
    bool globalValid = false;
    int64_t stride = 0;         ///< ELL column stride (>= end - begin, multiple of 64)
    int64_t totalFull = 0;
    /// full rows, column-major nbr[k * stride + r]; bit 31 marks the entries that belong
    /// to the half list (Cabana discriminator evaluated at build time)
    mrmd_hip::DBuf<uint32_t> nbr;
    mrmd_hip::DBuf<int32_t> cnt;  ///< full count per row
    bool halfValid = false;
    int32_t capHalf = 0;
    mrmd_hip::DBuf<int32_t> hnbr;  ///< half rows (Cabana discriminator), same layout
    mrmd_hip::DBuf<int32_t> hcnt;
    mrmd_hip::Binning bin;

    // ---- molecule lists whose every record is one atom (checked at build time) ----
    const mrmd_molecules* monoOf = nullptr;  ///< the molecules, when all records were monatomic
    uint64_t monoVersion = 0;                ///< their layout version at that build

    /// sorted slot of each binned record (inverse of bin.sorted), made on demand per build
    mrmd_hip::DBuf<int32_t> slotOf;
    bool slotOfValid = false;
    /// set by a caller that has itself written the current positions of every binned record
    /// into bin.sortedPos; the next force call then skips its gather and clears the flag
    bool sortedFresh = false;

    /// copy the current positions into sorted-slot order (bin.sortedPos) for the tiled kernels
    void refreshSorted(const double4* pos);
    /// the inverse: pos[record] = bin.sortedPos[slot] for every binned record (time loops that
    /// kept the positions in slot order alone)
    void positionsFromSlots(double4* pos);
    void positionsFromSlotsNoThrow(double4* pos) noexcept;
    /// slotOf, inverted from bin.sorted when the list was rebuilt since the last call
    const int32_t* ensureSlotOf();
    /// materialise the global-index full rows from the tiled list (no-op when current)
    void ensureGlobal();
    /// materialise the half list from the flagged full rows (no-op when current)
    void ensureHalf();
};
