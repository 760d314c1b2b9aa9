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
// order is ascending record index inside every cell (a stable counting sort)
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

    void build(mrmd_ctx* ctx, const double4* pos, int64_t b, int64_t e, const GridD& g, bool withPos);
};
}  // namespace mrmd_hip

struct mrmd_cell_list
{
    mrmd_hip::CtxPtr ctx;
    mrmd_hip::Binning bin;
};

struct mrmd_verlet_list
{
    mrmd_hip::CtxPtr ctx;
    int64_t nrows = 0;          ///< records of the binned slice (counts exist for all)
    int64_t begin = 0, end = 0; ///< rows = records [begin, end)
    int64_t stride = 0;         ///< ELL column stride (>= end - begin, multiple of 64)
    double rn = 0, ratio = 1;
    int range = 1;
    int32_t capFull = 0;
    int32_t maxFull = 0;
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

    /// materialise the half list from the flagged full rows (no-op when current)
    void ensureHalf();
};
