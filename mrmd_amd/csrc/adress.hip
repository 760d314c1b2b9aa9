// AdResS LJ / ideal-gas coupling on the device: slab weighting, molecule update, the
// weighted molecule-pair force with drift and compensation terms, molecule -> atom force
// distribution, and the thermodynamic force (density sampling, smoothing, gradient,
// application).
//
// LJ_IdealGas keeps the reference's half-list algorithm (LJ_IdealGas.cpp:52-225): the
// weighting function is evaluated on ghost images, whose lambda need not equal that of
// their real molecule, so the pair must be evaluated once with the ghost's lambda exactly
// as the reference does; partner forces go through FP64 atomics.  Only molecule pairs
// with at least one non-CG partner do pair work.  Density sampling accumulates into a
// workgroup-private LDS histogram before one global atomic per touched bin.
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

__global__ void k_sumOrdered(const double* partials, int64_t n, double* out)
{
    __shared__ double red[BLOCK / WAVE];
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) s += partials[k];
    s = blockSum(s, red);
    if (threadIdx.x == 0) *out = s;
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
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    {
        const double4 p = pos[i];
        const int64_t bin = getBin(p.x, h.min, h.invBinSize, h.numBins);
        if (bin == -1) continue;
        atomicAdd(&loc[bin * h.numHist + int64_t(p.w)], 1.0);
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

int mrmd_ljig_run(mrmd_lj_idealgas* s, mrmd_molecules* m, mrmd_verlet_list* vl, mrmd_atoms* a, double* energy)
{
    MRMD_API_BEGIN(s ? s->ctx : nullptr)
    REQUIRE(s && m && vl && a, "invalid arguments");
    REQUIRE(vl->begin == 0, "the molecule list rows must start at record 0");
    REQUIRE(a->maxType < s->numTypes, "atom type %lld outside the %lld-type table", (long long)a->maxType,
            (long long)s->numTypes);
    REQUIRE(vl->nrows <= m->nlocal + m->nghost, "the list was built for more molecules than exist");
    mrmd_ctx* ctx = s->ctx;
    vl->ensureHalf();
    const int64_t rows = std::min<int64_t>(m->nlocal, vl->end);
    const unsigned grid = gridFor(rows);
    double* part = static_cast<double*>(ctxScratch(ctx, 1, size_t(grid + 1) * sizeof(double)));
    const int64_t nh = COMP_BINS * s->numTypes;
    const double compBin = (1.0 - 0.0) / double(COMP_BINS);
    IGArgs g{m->view(), a->view(), vl->hnbr.p, vl->hcnt.p, vl->stride, rows, s->dtable.p, int(s->numTypes), s->rcSqr,
             (s->runCounter % s->samplingInterval) == 0 ? 1 : 0, 0.0, 1.0 / compBin, s->hist.p, s->hist.p + nh,
             s->hist.p + 2 * nh};
    if (rows > 0)
    {
        KTimer t(ctx, "lj_idealgas");
        hipLaunchKernelGGL(k_ljIdealGas, dim3(grid), dim3(BLOCK), 0, ctx->stream, g, part);
        hipLaunchKernelGGL(k_sumOrdered, dim3(1), dim3(BLOCK), 0, ctx->stream, part, int64_t(grid), part + grid);
    }
    else
        hipLaunchKernelGGL(k_zeroD, dim3(1), dim3(64), 0, ctx->stream, part + grid, int64_t(1));
    if (s->runCounter % s->updateInterval == 0)
        hipLaunchKernelGGL(k_updateMeanComp, dim3(gridFor(nh)), dim3(BLOCK), 0, ctx->stream, s->hist.p, s->hist.p + nh,
                           s->hist.p + 2 * nh, nh, 10.0);
    checkLaunch();
    ++s->runCounter;
    if (energy)
    {
        HIPCHECK(hipMemcpyAsync(ctx->pinned, part + grid, 8, hipMemcpyDeviceToHost, ctx->stream));
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

int mrmd_tf_sample(mrmd_thermo_force* t, const mrmd_atoms* a)
{
    MRMD_API_BEGIN(t ? t->ctx : nullptr)
    REQUIRE(t && a, "invalid arguments");
    REQUIRE(a->maxType < t->hd.numHist, "atom type outside the density profile");
    if (a->nlocal > 0)
    {
        KTimer k(t->ctx, "tf_sample");
        const unsigned grid = std::min<unsigned>(gridFor(a->nlocal), 1024u);
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
