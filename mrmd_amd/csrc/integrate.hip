// Velocity-Verlet and Langevin (BAOAB-style) integration and the kinetic energy.
// Streaming kernels: one lane per local atom, coalesced SoA reads/writes; the maximum
// displacement is reduced with wave shuffles to one value per workgroup and then by one
// small ordered pass (no atomics on a single address).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>

#include "containers.hpp"
#include "detmath.hpp"
#include "physics.hpp"

using namespace mrmd_hip;

namespace
{
// device scalars of an atoms container: [0] max |dx|^2 of the last pre-step
struct DispScalars
{
    double maxSq;
};

__global__ void k_zero1(double* p) { *p = 0.0; }

// VelocityVerlet::preForceIntegrate (VelocityVerlet.cpp:26-67; UpdateSteps.hpp:25-54)
__global__ __launch_bounds__(BLOCK) void k_vvPre(AtomsView a, int64_t n, double dt, double* partials)
{
    __shared__ double red[BLOCK / WAVE];
    double d2 = 0.0;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
    {
        const double dtHalf = 0.5 * dt;
        double4 p = a.pos[i];
        const double m = a.mass[i];
        double vx = a.vel[i], vy = a.vel[a.cap + i], vz = a.vel[2 * a.cap + i];
        const double fx = a.force[i], fy = a.force[a.cap + i], fz = a.force[2 * a.cap + i];
        const double dtfm = dtHalf / m;
        vx += dtfm * fx;
        vy += dtfm * fy;
        vz += dtfm * fz;
        const double ox = p.x, oy = p.y, oz = p.z;
        p.x += dt * vx;
        p.y += dt * vy;
        p.z += dt * vz;
        a.pos[i] = p;
        a.vel[i] = vx;
        a.vel[a.cap + i] = vy;
        a.vel[2 * a.cap + i] = vz;
        const double dx = ox - p.x, dy = oy - p.y, dz = oz - p.z;
        d2 = fmax(d2, dx * dx + dy * dy + dz * dz);
    }
    const double m = blockMax(d2, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

// VelocityVerlet::postForceIntegrate (VelocityVerlet.cpp:69-90)
__global__ __launch_bounds__(BLOCK) void k_vvPost(AtomsView a, int64_t n, double dt)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dtfm = (0.5 * dt) / a.mass[i];
    a.vel[i] += dtfm * a.force[i];
    a.vel[a.cap + i] += dtfm * a.force[a.cap + i];
    a.vel[2 * a.cap + i] += dtfm * a.force[2 * a.cap + i];
}

// VelocityVerletLangevinThermostat::preForceIntegrate_apply_if
// (VelocityVerletLangevinThermostat.hpp:63-136; UpdateSteps.hpp:56-79)
// mass-dependent factors of the O-step for one mass (the expressions of the per-atom path)
struct LangevinFactors
{
    double m, dtfm, damping, sigma;
};

__host__ __device__ inline LangevinFactors langevinFactors(double m, double dt, double zeta, double temperature)
{
    LangevinFactors f;
    f.m = m;
    f.dtfm = (0.5 * dt) / m;
    const double dtm = dt / m;
    f.damping = detExp(-zeta * dtm);
    f.sigma = sqrt(temperature / m * (1.0 - detExp(-2.0 * zeta * dtm)));
    return f;
}

// UNI: every atom has the mass uf.m, whose factors the host computed with the same code
template <bool PRED, bool UNI>
__global__ __launch_bounds__(BLOCK) void k_langevinPre(AtomsView a, int64_t n, double dt, double zeta, double temperature,
                                                       uint64_t seed, uint64_t call, mrmd_pred1 pred, LangevinFactors uf,
                                                       double* partials)
{
    __shared__ double red[BLOCK / WAVE];
    __shared__ WaveDraws draws[BLOCK / WAVE];
    double d2 = 0.0;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool live = i < n;
    const double dtHalf = 0.5 * dt;
    double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
    double vx = 0.0, vy = 0.0, vz = 0.0, ox = 0.0, oy = 0.0, oz = 0.0;
    LangevinFactors f = uf;
    if (live)
    {
        p = a.pos[i];
        vx = a.vel[i];
        vy = a.vel[a.cap + i];
        vz = a.vel[2 * a.cap + i];
        if (!UNI) f = langevinFactors(a.mass[i], dt, zeta, temperature);
        const double dtfm = f.dtfm;
        vx += dtfm * a.force[i];
        vy += dtfm * a.force[a.cap + i];
        vz += dtfm * a.force[2 * a.cap + i];
        ox = p.x;
        oy = p.y;
        oz = p.z;
        p.x += dtHalf * vx;
        p.y += dtHalf * vy;
        p.z += dtHalf * vz;
    }
    const bool need = live && (!PRED || evalPred1(pred, p.x, p.y, p.z));
    if (__ballot(need) != 0)
    {
        double xi[3];
        waveNormals3(seed, uint64_t(i), call, need, draws[threadIdx.x / WAVE], xi);
        if (need)
        {
            const double damping = f.damping;
            const double sigma = f.sigma;
            vx *= damping;
            vy *= damping;
            vz *= damping;
            vx += sigma * xi[0];
            vy += sigma * xi[1];
            vz += sigma * xi[2];
        }
    }
    if (live)
    {
        p.x += dtHalf * vx;
        p.y += dtHalf * vy;
        p.z += dtHalf * vz;
        a.pos[i] = p;
        a.vel[i] = vx;
        a.vel[a.cap + i] = vy;
        a.vel[2 * a.cap + i] = vz;
        const double dx = ox - p.x, dy = oy - p.y, dz = oz - p.z;
        d2 = fmax(d2, dx * dx + dy * dy + dz * dz);
    }
    const double mx = blockMax(d2, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = mx;
}

// A step's post-force integration (k_vvPost) fused with the next step's pre-force
// integration, for time loops that run both back to back: per atom the two kernels'
// arithmetic in the same order (the force is the same between them), so the results are
// bit-identical, with one pass over the arrays instead of two.  LANGEVIN: the pre-force
// part is k_langevinPre<false, UNI>, otherwise k_vvPre.  sortedPos (if not null): each new
// position also goes to the atom's sorted slot slotOf[i] of the Verlet list.
// ADR (the AdResS loop, molecule k = atom k): the atom's molecule force times its relative
// mass is added to F first (ContributeMoleculeForceToAtoms, ContributeMoleculeForceToAtoms.cpp:
// 23-48), then the thermodynamic force of its x-bin to F_x (ThermodynamicForce::apply,
// ThermodynamicForce.hpp:98-122), the sums those actions leave in the force array
struct AdrFold
{
    const double* molForce;  ///< [3][molCap]
    int64_t molCap;
    const double* tfForce;
    HistD hd;
};

// SLOTS: the positions live in sortedPos alone (the time loop's slot-order copy, current for
// every record): each atom's position is read from and stored to its slot only, and the
// positions array is left as it was
template <bool LANGEVIN, bool UNI, bool ADR = false, bool SLOTS = false>
__global__ __launch_bounds__(BLOCK) void k_postPre(AtomsView a, int64_t n, double dt, double zeta, double temperature,
                                                   uint64_t seed, uint64_t call, LangevinFactors uf, double* partials,
                                                   double4* sortedPos, const int32_t* slotOf, AdrFold adr)
{
    __shared__ double red[BLOCK / WAVE];
    __shared__ WaveDraws draws[LANGEVIN ? BLOCK / WAVE : 1];
    double d2 = 0.0;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool live = i < n;
    const double dtHalf = 0.5 * dt;
    double4 p = make_double4(0.0, 0.0, 0.0, 0.0);
    double vx = 0.0, vy = 0.0, vz = 0.0, ox = 0.0, oy = 0.0, oz = 0.0;
    int32_t slot = 0;
    LangevinFactors f = uf;
    if (live)
    {
        double fx = a.force[i], fy = a.force[a.cap + i], fz = a.force[2 * a.cap + i];
        if (ADR)
        {
            const double rm = a.relmass[i];
            fx += rm * adr.molForce[i];
            fy += rm * adr.molForce[adr.molCap + i];
            fz += rm * adr.molForce[2 * adr.molCap + i];
            const double4 q = a.pos[i];
            const int64_t bin = getBin(q.x, adr.hd.min, adr.hd.invBinSize, adr.hd.numBins);
            if (bin != -1) fx += adr.tfForce[bin * adr.hd.numHist + int64_t(q.w)];
        }
        vx = a.vel[i];
        vy = a.vel[a.cap + i];
        vz = a.vel[2 * a.cap + i];
        // post-force kick: (0.5 dt) / m, the expression of k_vvPost and of langevinFactors
        if (LANGEVIN && !UNI) f = langevinFactors(a.mass[i], dt, zeta, temperature);
        const double dtfm = UNI ? uf.dtfm : (LANGEVIN ? f.dtfm : dtHalf / a.mass[i]);
        vx += dtfm * fx;
        vy += dtfm * fy;
        vz += dtfm * fz;
        // pre-force kick and drift of the next step
        if (SLOTS)
        {
            slot = slotOf[i];
            p = sortedPos[slot];
        }
        else
            p = a.pos[i];
        vx += dtfm * fx;
        vy += dtfm * fy;
        vz += dtfm * fz;
        ox = p.x;
        oy = p.y;
        oz = p.z;
        if (LANGEVIN)
        {
            p.x += dtHalf * vx;
            p.y += dtHalf * vy;
            p.z += dtHalf * vz;
        }
        else
        {
            p.x += dt * vx;
            p.y += dt * vy;
            p.z += dt * vz;
        }
    }
    if (LANGEVIN && __ballot(live) != 0)
    {
        double xi[3];
        waveNormals3(seed, uint64_t(i), call, live, draws[LANGEVIN ? threadIdx.x / WAVE : 0], xi);
        if (live)
        {
            vx *= f.damping;
            vy *= f.damping;
            vz *= f.damping;
            vx += f.sigma * xi[0];
            vy += f.sigma * xi[1];
            vz += f.sigma * xi[2];
            p.x += dtHalf * vx;
            p.y += dtHalf * vy;
            p.z += dtHalf * vz;
        }
    }
    if (live)
    {
        if (SLOTS)
            sortedPos[slot] = p;
        else
        {
            a.pos[i] = p;
            if (sortedPos) sortedPos[slotOf[i]] = p;
        }
        a.vel[i] = vx;
        a.vel[a.cap + i] = vy;
        a.vel[2 * a.cap + i] = vz;
        const double dx = ox - p.x, dy = oy - p.y, dz = oz - p.z;
        d2 = fmax(d2, dx * dx + dy * dy + dz * dz);
    }
    const double mx = blockMax(d2, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = mx;
}

// analysis::getKineticEnergy (KineticEnergy.hpp:26-39): per-workgroup partials
__global__ __launch_bounds__(BLOCK) void k_kinetic(AtomsView a, int64_t n, double* partials)
{
    __shared__ double red[BLOCK / WAVE];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i < n)
    {
        const double vx = a.vel[i], vy = a.vel[a.cap + i], vz = a.vel[2 * a.cap + i];
        s = a.mass[i] * (vx * vx + vy * vy + vz * vz);
    }
    s = blockSum(s, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ void k_sumOrdered(const double* partials, int64_t n, double* out)
{
    __shared__ double red[BLOCK / WAVE];
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += blockDim.x) s += partials[k];
    s = blockSum(s, red);
    if (threadIdx.x == 0) *out = s;
}

// BerendsenThermostat::apply (BerendsenThermostat.cpp:27-50)
__global__ void k_scaleVel(AtomsView a, int64_t n, double beta)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a.vel[i] *= beta;
    a.vel[a.cap + i] *= beta;
    a.vel[2 * a.cap + i] *= beta;
}

// per-workgroup maxima of the squared displacement, then their maximum (one workgroup of
// 1,024 lanes, eight independent loads per lane in flight: a few rounds of loads even at
// millions of atoms, whose latency, not bandwidth, bounds this kernel)
// host (if not null) is coherent pinned host memory that receives the result as two tagged
// words (low and high half of the double, each with the call's tag in its high half): the
// host takes the value as soon as both words carry the tag, without waiting for the stream
constexpr int MAXP_THREADS = 1024;
__global__ __launch_bounds__(MAXP_THREADS) void k_maxPartials(const double* partials, int64_t n, double* out,
                                                              uint64_t* host, uint32_t tag)
{
    __shared__ double red[MAXP_THREADS / WAVE];
    double m[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int64_t k = threadIdx.x;
    for (; k + 7 * MAXP_THREADS < n; k += 8 * MAXP_THREADS)
    {
#pragma unroll
        for (int u = 0; u < 8; ++u) m[u] = fmax(m[u], partials[k + u * MAXP_THREADS]);
    }
    for (; k < n; k += MAXP_THREADS) m[0] = fmax(m[0], partials[k]);
    double r = fmax(fmax(fmax(m[0], m[1]), fmax(m[2], m[3])), fmax(fmax(m[4], m[5]), fmax(m[6], m[7])));
    r = blockMax(r, red);
    if (threadIdx.x == 0)
    {
        *out = r;
        if (host)
        {
            const uint64_t bits = uint64_t(__double_as_longlong(r));
            __hip_atomic_store(host, tagWord(tag, uint32_t(bits)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(host + 1, tagWord(tag, uint32_t(bits >> 32)), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// scratch slot 3: [0] maximum squared displacement of the last pre-step, [1 ..] partials
double* dispScalar(mrmd_atoms* a) { return static_cast<double*>(ctxScratch(a->ctx, 3, 64)); }
double* dispPartials(mrmd_atoms* a, unsigned grid)
{
    return static_cast<double*>(ctxScratch(a->ctx, 3, (size_t(grid) + 1) * sizeof(double))) + 1;
}

// The reduction kernel posts the maximum as the two tagged words at DISP_WORD of the pinned
// buffer (dispHost writes them itself when no kernel runs); the host polls for the call's tag
// (waitTags).
void finishDisp(mrmd_atoms* a, double* maxDisp, uint32_t tag)
{
    if (!maxDisp) return;
    mrmd_ctx* c = a->ctx;
    waitTags(c, DISP_WORD, 2, tag);
    const uint64_t bits = uint64_t(tagValue(c, DISP_WORD)) | (uint64_t(tagValue(c, DISP_WORD + 1)) << 32);
    double d2;
    std::memcpy(&d2, &bits, sizeof(d2));
    *maxDisp = sqrt(d2);
}

// target of the reduction's host copy (the context's pinned words) and the call's tag when
// the caller wants the value; nlocal = 0: no kernel writes them, so they are set here
uint64_t* dispHost(mrmd_atoms* a, const double* maxDisp, uint32_t* tag)
{
    *tag = 0;
    if (!maxDisp) return nullptr;
    mrmd_ctx* c = a->ctx;
    *tag = nextTag(c);
    uint64_t* w = reinterpret_cast<uint64_t*>(c->pinned) + DISP_WORD;
    if (a->nlocal == 0)
    {
        __atomic_store_n(w, tagWord(*tag, 0u), __ATOMIC_RELEASE);
        __atomic_store_n(w + 1, tagWord(*tag, 0u), __ATOMIC_RELEASE);
    }
    return w;
}
}  // namespace

// the fused post/pre step of time loops (k_postPre); synchronises for the displacement
int mrmd_hip::fusedPostPre(mrmd_atoms* a, double dt, bool langevin, double zeta, double temperature, uint64_t seed,
                           uint64_t call, double* maxDisp, double4* sortedPos, const int32_t* slotOf,
                           const mrmd_molecules* adrMol, const double* tfForce, const HistD* tfHd, bool slotsOnly)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && maxDisp, "invalid arguments");
    mrmd_ctx* ctx = a->ctx;
    const unsigned grid = gridFor(a->nlocal);
    double* part = dispPartials(a, grid);
    double* ms = dispScalar(a);
    uint32_t tag;
    uint64_t* host = dispHost(a, maxDisp, &tag);
    if (a->nlocal == 0) hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, ctx->stream, ms);
    if (a->nlocal > 0)
    {
        KTimer t(ctx, "post_pre");
        const bool uni = a->massUniform && a->massValue > 0.0;
        const LangevinFactors uf = uni ? langevinFactors(a->massValue, dt, zeta, temperature) : LangevinFactors{};
        const AdrFold adr{adrMol ? adrMol->force.p : nullptr, adrMol ? adrMol->cap : 0, tfForce,
                          tfHd ? *tfHd : HistD{}};
        auto launch = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(BLOCK), 0, ctx->stream, a->view(), a->nlocal, dt, zeta,
                               temperature, seed, call, uf, part, sortedPos, slotOf, adr);
        };
        if (adrMol)
        {
            REQUIRE(langevin && tfForce && tfHd && a->maxType < tfHd->numHist && adrMol->nlocal == a->nlocal,
                    "the fused AdResS step needs Langevin steps, a molecule per local atom and a type column per "
                    "atom type");
            uni ? launch(k_postPre<true, true, true>) : launch(k_postPre<true, false, true>);
        }
        else if (slotsOnly)
        {
            REQUIRE(sortedPos && slotOf, "positions in slot order need the slot copy and map");
            if (langevin)
                uni ? launch(k_postPre<true, true, false, true>) : launch(k_postPre<true, false, false, true>);
            else
                uni ? launch(k_postPre<false, true, false, true>) : launch(k_postPre<false, false, false, true>);
        }
        else if (langevin)
            uni ? launch(k_postPre<true, true>) : launch(k_postPre<true, false>);
        else
            uni ? launch(k_postPre<false, true>) : launch(k_postPre<false, false>);
        hipLaunchKernelGGL(k_maxPartials, dim3(1), dim3(MAXP_THREADS), 0, ctx->stream, part, int64_t(grid), ms, host,
                           tag);
    }
    checkLaunch();
    finishDisp(a, maxDisp, tag);
    MRMD_API_END
}

extern "C" {

int mrmd_vv_pre(mrmd_atoms* a, double dt, double* maxDisp)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a, "null atoms");
    mrmd_ctx* ctx = a->ctx;
    const unsigned grid = gridFor(a->nlocal);
    double* part = dispPartials(a, grid);
    double* ms = dispScalar(a);
    uint32_t tag;
    uint64_t* host = dispHost(a, maxDisp, &tag);
    if (a->nlocal > 0)
    {
        KTimer t(ctx, "vv_pre");
        hipLaunchKernelGGL(k_vvPre, dim3(grid), dim3(BLOCK), 0, ctx->stream, a->view(), a->nlocal, dt, part);
        hipLaunchKernelGGL(k_maxPartials, dim3(1), dim3(MAXP_THREADS), 0, ctx->stream, part, int64_t(grid), ms, host,
                           tag);
    }
    else
        hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, ctx->stream, ms);
    checkLaunch();
    finishDisp(a, maxDisp, tag);
    MRMD_API_END
}

int mrmd_vv_post(mrmd_atoms* a, double dt)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a, "null atoms");
    if (a->nlocal == 0) return MRMD_OK;
    KTimer t(a->ctx, "vv_post");
    hipLaunchKernelGGL(k_vvPost, dim3(gridFor(a->nlocal)), dim3(BLOCK), 0, a->ctx->stream, a->view(), a->nlocal, dt);
    checkLaunch();
    MRMD_API_END
}

int mrmd_langevin_pre(mrmd_atoms* a, double dt, double zeta, double temperature, uint64_t seed, uint64_t call,
                      const mrmd_pred1* pred, double* maxDisp)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a, "null atoms");
    validatePred1(pred);
    mrmd_ctx* ctx = a->ctx;
    const unsigned grid = gridFor(a->nlocal);
    double* part = dispPartials(a, grid);
    double* ms = dispScalar(a);
    uint32_t tag;
    uint64_t* host = dispHost(a, maxDisp, &tag);
    if (a->nlocal == 0) hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, ctx->stream, ms);
    if (a->nlocal > 0)
    {
        KTimer t(ctx, "langevin_pre");
        const bool usePred = pred && pred->kind != MRMD_PRED_ALWAYS;
        const mrmd_pred1 p = pred ? *pred : mrmd_pred1{};
        const bool uni = a->massUniform && a->massValue > 0.0;
        const LangevinFactors uf = uni ? langevinFactors(a->massValue, dt, zeta, temperature) : LangevinFactors{};
        auto launch = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(grid), dim3(BLOCK), 0, ctx->stream, a->view(), a->nlocal, dt, zeta,
                               temperature, seed, call, p, uf, part);
        };
        if (usePred)
            uni ? launch(k_langevinPre<true, true>) : launch(k_langevinPre<true, false>);
        else
            uni ? launch(k_langevinPre<false, true>) : launch(k_langevinPre<false, false>);
        hipLaunchKernelGGL(k_maxPartials, dim3(1), dim3(MAXP_THREADS), 0, ctx->stream, part, int64_t(grid), ms, host,
                           tag);
    }
    checkLaunch();
    finishDisp(a, maxDisp, tag);
    MRMD_API_END
}

int mrmd_berendsen_apply(mrmd_atoms* a, double current, double target, double gamma)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a, "null atoms");
    if (current <= 0.0) return MRMD_OK;
    REQUIRE(target > 0.0, "the target temperature must be positive");
    const double beta = sqrt(1.0 + gamma * (target / current - 1.0));
    if (a->nlocal > 0)
        hipLaunchKernelGGL(k_scaleVel, dim3(gridFor(a->nlocal)), dim3(BLOCK), 0, a->ctx->stream, a->view(), a->nlocal,
                           beta);
    checkLaunch();
    MRMD_API_END
}

int mrmd_kinetic_energy(const mrmd_atoms* ac, double* energy)
{
    auto* a = const_cast<mrmd_atoms*>(ac);
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && energy, "invalid arguments");
    mrmd_ctx* ctx = a->ctx;
    const unsigned g = gridFor(a->nlocal);
    double* part = static_cast<double*>(ctxScratch(ctx, 1, size_t(g + 1) * sizeof(double)));
    if (a->nlocal > 0)
        hipLaunchKernelGGL(k_kinetic, dim3(g), dim3(BLOCK), 0, ctx->stream, a->view(), a->nlocal, part);
    hipLaunchKernelGGL(k_sumOrdered, dim3(1), dim3(BLOCK), 0, ctx->stream, part, a->nlocal > 0 ? int64_t(g) : 0,
                       part + g);
    checkLaunch();
    HIPCHECK(hipMemcpyAsync(ctx->pinned, part + g, 8, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    *energy = 0.5 * ctx->pinned[0];
    MRMD_API_END
}

}  // extern "C"
