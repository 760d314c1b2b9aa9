// Context, containers and device scans.
#include <algorithm>
#include <cstring>

#include <atomic>
#include <chrono>

#include "containers.hpp"

using namespace mrmd_hip;

// =========================================================================================
// device scans: exclusive prefix sums over int32 counts, tiles of 1024 (256 threads x 4)
namespace
{
constexpr int TILE = BLOCK * 4;

template <class TO>
__global__ void k_tileSums(const int32_t* in, int64_t n, TO* sums)
{
    __shared__ TO lds[BLOCK / WAVE];
    const int64_t base = int64_t(blockIdx.x) * TILE + int64_t(threadIdx.x) * 4;
    TO s = 0;
    for (int k = 0; k < 4; ++k)
        if (base + k < n) s += TO(in[base + k]);
    TO tot;
    blockExclusiveScan<TO>(s, lds, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// single block: exclusive scan of sums[0, m) in place, grand total to *total
template <class TO>
__global__ void k_scanSums(TO* sums, int64_t m, TO* total)
{
    __shared__ TO lds[BLOCK / WAVE];
    TO carry = 0;
    for (int64_t base = 0; base < m; base += BLOCK)
    {
        const int64_t i = base + threadIdx.x;
        TO v = i < m ? sums[i] : TO(0);
        TO tot;
        TO ex = blockExclusiveScan<TO>(v, lds, &tot);
        if (i < m) sums[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

template <class TO>
__global__ void k_tileScan(const int32_t* in, int64_t n, const TO* sums, TO* out)
{
    __shared__ TO lds[BLOCK / WAVE];
    const int64_t base = int64_t(blockIdx.x) * TILE + int64_t(threadIdx.x) * 4;
    TO v[4];
    TO s = 0;
    for (int k = 0; k < 4; ++k)
    {
        v[k] = (base + k < n) ? TO(in[base + k]) : TO(0);
        s += v[k];
    }
    TO tot;
    TO ex = blockExclusiveScan<TO>(s, lds, &tot) + sums[blockIdx.x];
    for (int k = 0; k < 4; ++k)
        if (base + k < n)
        {
            out[base + k] = ex;
            ex += v[k];
        }
}

// one workgroup per array (array k at in + k * n): exclusive scan in runs of TILE with a
// carried sum, for arrays short enough that one launch beats the three of the tiled form
constexpr int64_t SMALL_SCAN = 4 * TILE;
template <class TO>
__global__ __launch_bounds__(BLOCK) void k_scanSmall(const int32_t* in, TO* out, int64_t n, TO* total)
{
    __shared__ TO lds[BLOCK / WAVE];
    in += int64_t(blockIdx.x) * n;
    out += int64_t(blockIdx.x) * n;
    TO carry = 0;
    for (int64_t t0 = 0; t0 < n; t0 += TILE)
    {
        const int64_t base = t0 + int64_t(threadIdx.x) * 4;
        TO v[4];
        TO s = 0;
        for (int k = 0; k < 4; ++k)
        {
            v[k] = (base + k < n) ? TO(in[base + k]) : TO(0);
            s += v[k];
        }
        TO tot;
        TO ex = blockExclusiveScan<TO>(s, lds, &tot) + carry;
        for (int k = 0; k < 4; ++k)
            if (base + k < n)
            {
                out[base + k] = ex;
                ex += v[k];
            }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) total[blockIdx.x] = carry;
}

template <class TO>
void scanImpl(mrmd_ctx* ctx, const int32_t* in, TO* out, int64_t n, TO* totalDev)
{
    if (n <= SMALL_SCAN)
    {
        hipLaunchKernelGGL(k_scanSmall<TO>, dim3(1), dim3(BLOCK), 0, ctx->stream, in, out, n, totalDev);
        checkLaunch();
        return;
    }
    const int64_t tiles = std::max<int64_t>(1, (n + TILE - 1) / TILE);
    TO* sums = static_cast<TO*>(ctxScratch(ctx, 0, size_t(tiles) * sizeof(TO)));
    hipLaunchKernelGGL(k_tileSums<TO>, dim3(unsigned(tiles)), dim3(BLOCK), 0, ctx->stream, in, n, sums);
    hipLaunchKernelGGL(k_scanSums<TO>, dim3(1), dim3(BLOCK), 0, ctx->stream, sums, tiles, totalDev);
    hipLaunchKernelGGL(k_tileScan<TO>, dim3(unsigned(tiles)), dim3(BLOCK), 0, ctx->stream, in, n, sums, out);
    checkLaunch();
}
}  // namespace

namespace mrmd_hip
{
void scanExclusiveI64(mrmd_ctx* ctx, const int32_t* in, int64_t* out, int64_t n, int64_t* totalDev)
{
    scanImpl<int64_t>(ctx, in, out, n, totalDev);
}
void scanExclusiveI32(mrmd_ctx* ctx, const int32_t* in, int32_t* out, int64_t n, int32_t* totalDev)
{
    scanImpl<int32_t>(ctx, in, out, n, totalDev);
}
void scanExclusiveI32Multi(mrmd_ctx* ctx, const int32_t* in, int32_t* out, int64_t n, int count, int32_t* totalsDev)
{
    // one workgroup per array, looping over runs of TILE: one launch instead of three per
    // array while the runs stay few
    if (n <= (count > 1 ? 16 : 1) * SMALL_SCAN)
    {
        hipLaunchKernelGGL(k_scanSmall<int32_t>, dim3(unsigned(count)), dim3(BLOCK), 0, ctx->stream, in, out, n,
                           totalsDev);
        checkLaunch();
        return;
    }
    for (int k = 0; k < count; ++k) scanImpl<int32_t>(ctx, in + k * n, out + k * n, n, totalsDev + k);
}

uint32_t nextTag(mrmd_ctx* ctx)
{
    if (++ctx->hostTag == 0)
    {
        // the counter wrapped: no word may still hold a tag that is about to be reused
        sync(ctx);
        std::memset(static_cast<void*>(ctx->pinned), 0, PINNED_WORDS * sizeof(uint64_t));
        ctx->hostTag = 1;
    }
    return ctx->hostTag;
}

void waitTags(mrmd_ctx* ctx, int word, int n, uint32_t tag)
{
    const uint64_t* w = reinterpret_cast<const uint64_t*>(ctx->pinned) + word;
    auto tagged = [&] {
        for (int k = 0; k < n; ++k)
            if (uint32_t(__atomic_load_n(w + k, __ATOMIC_ACQUIRE) >> 32) != tag) return false;
        return true;
    };
    const auto t0 = std::chrono::steady_clock::now();
    while (!tagged())
    {
        if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20))
        {
            sync(ctx);
            break;
        }
    }
    if (!tagged()) fail(MRMD_ERR_STATE, "a device result did not reach the host");
}
}  // namespace mrmd_hip

namespace
{
// n int32 values as tagged words into the coherent pinned buffer: agent-scope loads (the
// values of the previous kernels, from any XCD), one system-scope 64-bit store per word
__global__ void k_postI32(const int32_t* src, int n, uint64_t* host, uint32_t tag)
{
    const int k = threadIdx.x;
    if (k >= n) return;
    const int32_t v = __hip_atomic_load(src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(host + k, tagWord(tag, uint32_t(v)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
}  // namespace

namespace mrmd_hip
{
void readbackI32(mrmd_ctx* ctx, const int32_t* dev, int n, int32_t* out)
{
    if (n < 0 || n > READBACK_INTS) fail(MRMD_ERR_ARG, "readback of %d values (at most %d)", n, READBACK_INTS);
    if (n == 0) return;
    const uint32_t tag = nextTag(ctx);
    hipLaunchKernelGGL(k_postI32, dim3(1), dim3(64), 0, ctx->stream, dev, n,
                       reinterpret_cast<uint64_t*>(ctx->pinned) + READBACK_WORD, tag);
    checkLaunch();
    waitTags(ctx, READBACK_WORD, n, tag);
    for (int k = 0; k < n; ++k) out[k] = int32_t(tagValue(ctx, READBACK_WORD + k));
}
}  // namespace mrmd_hip

// =========================================================================================
// containers
namespace
{
__global__ void k_zeroAtoms(AtomsView a, int64_t begin, int64_t end)
{
    const int64_t i = begin + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= end) return;
    a.pos[i] = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int d = 0; d < 3; ++d)
    {
        a.vel[d * a.cap + i] = 0.0;
        a.force[d * a.cap + i] = 0.0;
    }
    a.mass[i] = 0.0;
    a.charge[i] = 0.0;
    a.relmass[i] = 0.0;
}

__global__ void k_zeroMols(MolView m, int64_t begin, int64_t end)
{
    const int64_t i = begin + int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= end) return;
    m.pos[i] = make_double4(0.0, 0.0, 0.0, 0.0);
    for (int d = 0; d < 3; ++d)
    {
        m.force[d * m.cap + i] = 0.0;
        m.grad[d * m.cap + i] = 0.0;
    }
    m.lambda[i] = 0.0;
    m.modLambda[i] = 0.0;
    m.offset[i] = 0;
    m.numAtoms[i] = 0;
}

// staging (row-major host layout) -> device layout
__global__ void k_putVec4(double4* p, const double* s, int64_t off, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 v = p[off + k];
    v.x = s[3 * k];
    v.y = s[3 * k + 1];
    v.z = s[3 * k + 2];
    p[off + k] = v;
}
__global__ void k_getVec4(const double4* p, double* s, int64_t off, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 v = p[off + k];
    s[3 * k] = v.x;
    s[3 * k + 1] = v.y;
    s[3 * k + 2] = v.z;
}
__global__ void k_putType(double4* p, const int64_t* s, int64_t off, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    p[off + k].w = double(s[k]);
}
__global__ void k_getType(const double4* p, int64_t* s, int64_t off, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    s[k] = int64_t(p[off + k].w);
}
__global__ void k_putSoA3(double* f, int64_t cap, const double* s, int64_t off, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    for (int d = 0; d < 3; ++d) f[d * cap + off + k] = s[3 * k + d];
}
__global__ void k_getSoA3(const double* f, int64_t cap, double* s, int64_t off, int64_t n)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    for (int d = 0; d < 3; ++d) s[3 * k + d] = f[d * cap + off + k];
}
__global__ void k_fill(double* p, int64_t n, double v)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}
// the three components of a vector field in one launch: component blockIdx.y at p + y * stride
__global__ void k_fill3(double* p, int64_t stride, int64_t n, double v)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[blockIdx.y * stride + k] = v;
}
__global__ void k_fillPos(double4* p, int64_t n, double v)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double4 q = p[k];
    q.x = v;
    q.y = v;
    q.z = v;
    p[k] = q;
}
__global__ void k_fillI64(int64_t* p, int64_t n, int64_t v)
{
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

template <class T>
void copyD2D(mrmd_ctx* ctx, T* dst, const T* src, int64_t n)
{
    if (n > 0) HIPCHECK(hipMemcpyAsync(dst, src, size_t(n) * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream));
}

// reallocates a component-major [3][cap] array to a new capacity
void regrow3(mrmd_ctx* ctx, DBuf<double>& b, int64_t oldCap, int64_t newCap, int64_t keep)
{
    DBuf<double> nb;
    nb.reserve(ctx, 3 * newCap);
    if (b.p)
        for (int d = 0; d < 3; ++d) copyD2D(ctx, nb.p + d * newCap, b.p + d * oldCap, keep);
    sync(ctx);
    b.release();
    b.p = nb.p;
    b.n = nb.n;
    b.ctx = ctx;
    nb.p = nullptr;
    nb.n = 0;
}

template <class T>
void regrow1(mrmd_ctx* ctx, DBuf<T>& b, int64_t newCap, int64_t keep)
{
    b.reserve(ctx, newCap, keep);
}
}  // namespace

void mrmd_atoms::ensureCap(int64_t need)
{
    if (need <= cap) return;
    const int64_t nc = growCap(need);
    regrow1(ctx, pos, nc, size);
    regrow3(ctx, vel, cap, nc, size);
    regrow3(ctx, force, cap, nc, size);
    regrow1(ctx, mass, nc, size);
    regrow1(ctx, charge, nc, size);
    regrow1(ctx, relmass, nc, size);
    cap = nc;
}

void mrmd_atoms::setSize(int64_t n)
{
    ensureCap(n);
    if (n > size)
    {
        hipLaunchKernelGGL(k_zeroAtoms, dim3(gridFor(n - size)), dim3(BLOCK), 0, ctx->stream, view(), size, n);
        checkLaunch();
        if (massValue != 0.0) massUniform = false;
        if (chargeValue != 0.0) chargeUniform = false;
        if (relmassValue != 0.0) relmassUniform = false;
    }
    size = n;
}

void mrmd_molecules::ensureCap(int64_t need)
{
    if (need <= cap) return;
    const int64_t nc = growCap(need);
    regrow1(ctx, pos, nc, size);
    regrow3(ctx, force, cap, nc, size);
    regrow1(ctx, lambda, nc, size);
    regrow1(ctx, modLambda, nc, size);
    regrow3(ctx, grad, cap, nc, size);
    regrow1(ctx, offset, nc, size);
    regrow1(ctx, numAtoms, nc, size);
    cap = nc;
}

void mrmd_molecules::setSize(int64_t n)
{
    ++layoutVersion;
    ensureCap(n);
    if (n > size)
    {
        hipLaunchKernelGGL(k_zeroMols, dim3(gridFor(n - size)), dim3(BLOCK), 0, ctx->stream, view(), size, n);
        checkLaunch();
    }
    size = n;
}

extern "C" {

// ---------------------------------------------------------------------------------------
int mrmd_ctx_create(int device, mrmd_ctx** out)
{
    if (!out) return MRMD_ERR_ARG;
    *out = nullptr;
    auto* c = new mrmd_ctx();
    MRMD_API_BEGIN(c)
    int n = 0;
    HIPCHECK(hipGetDeviceCount(&n));
    REQUIRE(device >= 0 && device < n, "device %d not available (%d devices)", device, n);
    c->device = device;
    HIPCHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        fail(MRMD_ERR_UNSUP, "this library is built for gfx950, device is %s", prop.gcnArchName);
    HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->ownStream = true;
    ljNoAllocSetup(device);
    // coherent: kernels write scalars here for the host to read while the stream runs on
    HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&c->pinned), 64 * sizeof(double), hipHostMallocCoherent));
    std::memset(static_cast<void*>(c->pinned), 0, PINNED_WORDS * sizeof(uint64_t));
    *out = c;
    return MRMD_OK;
    }
    catch (const ::mrmd_hip::Error& e)
    {
        fprintf(stderr, "mrmd_ctx_create: %s\n", e.what());
        delete c;
        return e.code;
    }
    catch (...)
    {
        delete c;
        return MRMD_ERR_HIP;
    }
}

int mrmd_ctx_destroy(mrmd_ctx* c)
{
    if (!c) return MRMD_OK;
    ctxRelease(c);
    return MRMD_OK;
}
}  // extern "C"

void mrmd_hip::ctxRetain(mrmd_ctx* c) { ++c->refs; }

void mrmd_hip::ctxRelease(mrmd_ctx* c)
{
    if (--c->refs > 0) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->stats)
        for (auto& pr : kv.second.pending)
        {
            (void)hipEventDestroy(pr.first);
            (void)hipEventDestroy(pr.second);
        }
    for (auto e : c->eventPool) (void)hipEventDestroy(e);
    if (c->pinned) (void)hipHostFree(c->pinned);
    for (int s = 0; s < 4; ++s)
        if (c->scratch[s]) (void)hipFree(c->scratch[s]);
    if (c->ownStream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" {

const char* mrmd_ctx_last_error(const mrmd_ctx* c) { return c ? c->err.c_str() : "null context"; }

int mrmd_ctx_set_stream(mrmd_ctx* c, void* s)
{
    MRMD_API_BEGIN(c)
    REQUIRE(c, "null context");
    HIPCHECK(hipStreamSynchronize(c->stream));
    if (s)
    {
        if (c->ownStream) HIPCHECK(hipStreamDestroy(c->stream));
        c->stream = static_cast<hipStream_t>(s);
        c->ownStream = false;
    }
    else if (!c->ownStream)
    {
        HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->ownStream = true;
    }
    MRMD_API_END
}

int mrmd_ctx_synchronize(mrmd_ctx* c)
{
    MRMD_API_BEGIN(c)
    sync(c);
    MRMD_API_END
}

int mrmd_ctx_set_profiling(mrmd_ctx* c, int enable)
{
    MRMD_API_BEGIN(c)
    c->profiling = enable != 0;
    MRMD_API_END
}

int mrmd_ctx_set_profile_filter(mrmd_ctx* c, const char* kernel)
{
    MRMD_API_BEGIN(c)
    c->profileOnly = kernel ? kernel : "";
    MRMD_API_END
}

int mrmd_ctx_kernel_stats(mrmd_ctx* c, const char* kernel, int64_t* launches, double* total_ms)
{
    MRMD_API_BEGIN(c)
    flushStats(c);
    auto it = c->stats.find(kernel ? kernel : "");
    if (launches) *launches = it == c->stats.end() ? 0 : it->second.launches;
    if (total_ms) *total_ms = it == c->stats.end() ? 0.0 : it->second.ms;
    MRMD_API_END
}

int mrmd_ctx_reset_stats(mrmd_ctx* c)
{
    MRMD_API_BEGIN(c)
    flushStats(c);
    for (auto& kv : c->stats)
    {
        kv.second.launches = 0;
        kv.second.ms = 0.0;
    }
    MRMD_API_END
}

int mrmd_ctx_device_bytes(mrmd_ctx* c, int64_t* bytes)
{
    MRMD_API_BEGIN(c)
    if (bytes) *bytes = c->bytes;
    MRMD_API_END
}

// ---------------------------------------------------------------------------------------
int mrmd_atoms_create(mrmd_ctx* c, int64_t size, mrmd_atoms** out)
{
    MRMD_API_BEGIN(c)
    REQUIRE(out && size >= 0, "invalid arguments");
    auto* a = new mrmd_atoms();
    a->ctx = c;
    try
    {
        a->setSize(size);
    }
    catch (...)
    {
        delete a;
        throw;
    }
    *out = a;
    MRMD_API_END
}

int mrmd_atoms_destroy(mrmd_atoms* a)
{
    if (!a) return MRMD_OK;
    (void)hipStreamSynchronize(a->ctx->stream);
    delete a;
    return MRMD_OK;
}

int mrmd_atoms_resize(mrmd_atoms* a, int64_t size)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && size >= 0, "invalid arguments");
    a->setSize(size);
    MRMD_API_END
}

int mrmd_atoms_get_counts(const mrmd_atoms* a, int64_t* nl, int64_t* ng, int64_t* size)
{
    if (!a) return MRMD_ERR_ARG;
    if (nl) *nl = a->nlocal;
    if (ng) *ng = a->nghost;
    if (size) *size = a->size;
    return MRMD_OK;
}

int mrmd_atoms_set_counts(mrmd_atoms* a, int64_t nl, int64_t ng)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && nl >= 0 && ng >= 0, "invalid arguments");
    if (nl + ng > a->size) a->setSize(nl + ng);
    a->nlocal = nl;
    a->nghost = ng;
    MRMD_API_END
}

int mrmd_atoms_upload(mrmd_atoms* a, int field, const void* host, int64_t off, int64_t n)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && host && off >= 0 && n >= 0 && field >= 0 && field <= MRMD_RELMASS, "invalid arguments");
    if (n == 0) return MRMD_OK;
    if (field == MRMD_TYPE)
    {
        const int64_t* t = static_cast<const int64_t*>(host);
        for (int64_t k = 0; k < n; ++k)
        {
            REQUIRE(t[k] >= 0 && t[k] < 64, "atom type %lld outside [0, 64)", (long long)t[k]);
            a->maxType = std::max(a->maxType, t[k]);
        }
    }
    if (off + n > a->size) a->setSize(off + n);
    mrmd_ctx* ctx = a->ctx;
    const bool vec = field <= MRMD_FORCE;
    const int64_t bytes = (vec ? 3 : 1) * n * 8;
    a->stage.ensure(ctx, (vec ? 3 : 1) * n);
    HIPCHECK(hipMemcpyAsync(a->stage.p, host, size_t(bytes), hipMemcpyHostToDevice, ctx->stream));
    const dim3 g(gridFor(n)), b(BLOCK);
    switch (field)
    {
        case MRMD_POS:
            hipLaunchKernelGGL(k_putVec4, g, b, 0, ctx->stream, a->pos.p, a->stage.p, off, n);
            break;
        case MRMD_VEL:
            hipLaunchKernelGGL(k_putSoA3, g, b, 0, ctx->stream, a->vel.p, a->cap, a->stage.p, off, n);
            break;
        case MRMD_FORCE:
            hipLaunchKernelGGL(k_putSoA3, g, b, 0, ctx->stream, a->force.p, a->cap, a->stage.p, off, n);
            break;
        case MRMD_TYPE:
            hipLaunchKernelGGL(k_putType, g, b, 0, ctx->stream, a->pos.p,
                               reinterpret_cast<const int64_t*>(a->stage.p), off, n);
            break;
        case MRMD_MASS:
        case MRMD_CHARGE:
        case MRMD_RELMASS:
        {
            double* dst = field == MRMD_MASS ? a->mass.p : (field == MRMD_CHARGE ? a->charge.p : a->relmass.p);
            bool& uniform = field == MRMD_MASS ? a->massUniform
                                               : (field == MRMD_CHARGE ? a->chargeUniform : a->relmassUniform);
            double& value = field == MRMD_MASS ? a->massValue
                                               : (field == MRMD_CHARGE ? a->chargeValue : a->relmassValue);
            HIPCHECK(hipMemcpyAsync(dst + off, a->stage.p, size_t(n) * 8, hipMemcpyDeviceToDevice, ctx->stream));
            const double* m = static_cast<const double*>(host);
            bool same = true;
            for (int64_t k = 1; k < n && same; ++k) same = m[k] == m[0];
            if (n == 0)
                break;
            if (off == 0 && n >= a->size)
            {
                uniform = same;
                value = m[0];
            }
            else
                uniform = uniform && same && m[0] == value;
            break;
        }
    }
    checkLaunch();
    sync(ctx);
    MRMD_API_END
}

int mrmd_atoms_download(const mrmd_atoms* ac, int field, void* host, int64_t off, int64_t n)
{
    auto* a = const_cast<mrmd_atoms*>(ac);
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && host && off >= 0 && n >= 0 && off + n <= a->size && field >= 0 && field <= MRMD_RELMASS,
            "invalid arguments (offset %lld, n %lld, size %lld)", (long long)off, (long long)n,
            a ? (long long)a->size : -1LL);
    if (n == 0) return MRMD_OK;
    mrmd_ctx* ctx = a->ctx;
    const bool vec = field <= MRMD_FORCE;
    a->stage.ensure(ctx, (vec ? 3 : 1) * n);
    const dim3 g(gridFor(n)), b(BLOCK);
    const double* src = nullptr;
    switch (field)
    {
        case MRMD_POS:
            hipLaunchKernelGGL(k_getVec4, g, b, 0, ctx->stream, a->pos.p, a->stage.p, off, n);
            break;
        case MRMD_VEL:
            hipLaunchKernelGGL(k_getSoA3, g, b, 0, ctx->stream, a->vel.p, a->cap, a->stage.p, off, n);
            break;
        case MRMD_FORCE:
            hipLaunchKernelGGL(k_getSoA3, g, b, 0, ctx->stream, a->force.p, a->cap, a->stage.p, off, n);
            break;
        case MRMD_TYPE:
            hipLaunchKernelGGL(k_getType, g, b, 0, ctx->stream, a->pos.p, reinterpret_cast<int64_t*>(a->stage.p),
                               off, n);
            break;
        case MRMD_MASS:
            src = a->mass.p + off;
            break;
        case MRMD_CHARGE:
            src = a->charge.p + off;
            break;
        case MRMD_RELMASS:
            src = a->relmass.p + off;
            break;
    }
    checkLaunch();
    if (src)
        HIPCHECK(hipMemcpyAsync(host, src, size_t(n) * 8, hipMemcpyDeviceToHost, ctx->stream));
    else
        HIPCHECK(hipMemcpyAsync(host, a->stage.p, size_t((vec ? 3 : 1) * n) * 8, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    MRMD_API_END
}

int mrmd_atoms_fill(mrmd_atoms* a, int field, double value)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && field >= 0 && field <= MRMD_RELMASS, "invalid arguments");
    mrmd_ctx* ctx = a->ctx;
    const int64_t n = a->size;
    if (n == 0) return MRMD_OK;
    switch (field)
    {
        case MRMD_POS:
            hipLaunchKernelGGL(k_fillPos, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, a->pos.p, n, value);
            break;
        case MRMD_VEL:
        case MRMD_FORCE:
        {
            double* p = field == MRMD_VEL ? a->vel.p : a->force.p;
            hipLaunchKernelGGL(k_fill3, dim3(gridFor(n), 3), dim3(BLOCK), 0, ctx->stream, p, a->cap, n, value);
            break;
        }
        case MRMD_TYPE:
        {
            std::vector<int64_t> h(static_cast<size_t>(n), static_cast<int64_t>(value));
            return mrmd_atoms_upload(a, MRMD_TYPE, h.data(), 0, n);
        }
        case MRMD_MASS:
            hipLaunchKernelGGL(k_fill, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, a->mass.p, n, value);
            a->massUniform = true;
            a->massValue = value;
            break;
        case MRMD_CHARGE:
            hipLaunchKernelGGL(k_fill, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, a->charge.p, n, value);
            a->chargeUniform = true;
            a->chargeValue = value;
            break;
        case MRMD_RELMASS:
            hipLaunchKernelGGL(k_fill, dim3(gridFor(n)), dim3(BLOCK), 0, ctx->stream, a->relmass.p, n, value);
            a->relmassUniform = true;
            a->relmassValue = value;
            break;
    }
    checkLaunch();
    MRMD_API_END
}

int mrmd_atoms_device_ptr(mrmd_atoms* a, int field, void** ptr, int64_t* stride)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && ptr, "invalid arguments");
    int64_t s = 0;
    void* p = nullptr;
    switch (field)
    {
        case MRMD_POS:
            p = a->pos.p;
            break;
        case MRMD_VEL:
            p = a->vel.p;
            s = a->cap;
            break;
        case MRMD_FORCE:
            p = a->force.p;
            s = a->cap;
            break;
        case MRMD_MASS:
            p = a->mass.p;
            a->massUniform = false;  // the caller may write masses directly
            break;
        case MRMD_CHARGE:
            p = a->charge.p;
            a->chargeUniform = false;
            break;
        case MRMD_RELMASS:
            p = a->relmass.p;
            a->relmassUniform = false;
            break;
        default:
            fail(MRMD_ERR_ARG, "field %d has no separate device array", field);
    }
    *ptr = p;
    if (stride) *stride = s;
    MRMD_API_END
}

// ---------------------------------------------------------------------------------------
int mrmd_molecules_create(mrmd_ctx* c, int64_t size, mrmd_molecules** out)
{
    MRMD_API_BEGIN(c)
    REQUIRE(out && size >= 0, "invalid arguments");
    auto* m = new mrmd_molecules();
    m->ctx = c;
    try
    {
        m->setSize(size);
    }
    catch (...)
    {
        delete m;
        throw;
    }
    *out = m;
    MRMD_API_END
}

int mrmd_molecules_destroy(mrmd_molecules* m)
{
    if (!m) return MRMD_OK;
    (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
    return MRMD_OK;
}

int mrmd_molecules_resize(mrmd_molecules* m, int64_t size)
{
    MRMD_API_BEGIN(m ? m->ctx : nullptr)
    REQUIRE(m && size >= 0, "invalid arguments");
    m->setSize(size);
    MRMD_API_END
}

int mrmd_molecules_get_counts(const mrmd_molecules* m, int64_t* nl, int64_t* ng, int64_t* size)
{
    if (!m) return MRMD_ERR_ARG;
    if (nl) *nl = m->nlocal;
    if (ng) *ng = m->nghost;
    if (size) *size = m->size;
    return MRMD_OK;
}

int mrmd_molecules_set_counts(mrmd_molecules* m, int64_t nl, int64_t ng)
{
    MRMD_API_BEGIN(m ? m->ctx : nullptr)
    REQUIRE(m && nl >= 0 && ng >= 0, "invalid arguments");
    if (nl + ng > m->size) m->setSize(nl + ng);
    m->nlocal = nl;
    m->nghost = ng;
    ++m->layoutVersion;
    MRMD_API_END
}

int mrmd_molecules_upload(mrmd_molecules* m, int field, const void* host, int64_t off, int64_t n)
{
    MRMD_API_BEGIN(m ? m->ctx : nullptr)
    REQUIRE(m && host && off >= 0 && n >= 0 && field >= 0 && field <= MRMD_MOL_NUMATOMS, "invalid arguments");
    if (n == 0) return MRMD_OK;
    if (off + n > m->size) m->setSize(off + n);
    if (field == MRMD_MOL_OFFSET || field == MRMD_MOL_NUMATOMS) ++m->layoutVersion;
    mrmd_ctx* ctx = m->ctx;
    const bool vec = field == MRMD_MOL_POS || field == MRMD_MOL_FORCE || field == MRMD_MOL_GRADLAMBDA;
    m->stage.ensure(ctx, (vec ? 3 : 1) * n);
    HIPCHECK(hipMemcpyAsync(m->stage.p, host, size_t((vec ? 3 : 1) * n) * 8, hipMemcpyHostToDevice, ctx->stream));
    const dim3 g(gridFor(n)), b(BLOCK);
    switch (field)
    {
        case MRMD_MOL_POS:
            hipLaunchKernelGGL(k_putVec4, g, b, 0, ctx->stream, m->pos.p, m->stage.p, off, n);
            break;
        case MRMD_MOL_FORCE:
            hipLaunchKernelGGL(k_putSoA3, g, b, 0, ctx->stream, m->force.p, m->cap, m->stage.p, off, n);
            break;
        case MRMD_MOL_GRADLAMBDA:
            hipLaunchKernelGGL(k_putSoA3, g, b, 0, ctx->stream, m->grad.p, m->cap, m->stage.p, off, n);
            break;
        case MRMD_MOL_LAMBDA:
            HIPCHECK(hipMemcpyAsync(m->lambda.p + off, m->stage.p, size_t(n) * 8, hipMemcpyDeviceToDevice, ctx->stream));
            break;
        case MRMD_MOL_MODLAMBDA:
            HIPCHECK(hipMemcpyAsync(m->modLambda.p + off, m->stage.p, size_t(n) * 8, hipMemcpyDeviceToDevice, ctx->stream));
            break;
        case MRMD_MOL_OFFSET:
            HIPCHECK(hipMemcpyAsync(m->offset.p + off, m->stage.p, size_t(n) * 8, hipMemcpyDeviceToDevice, ctx->stream));
            break;
        case MRMD_MOL_NUMATOMS:
            HIPCHECK(hipMemcpyAsync(m->numAtoms.p + off, m->stage.p, size_t(n) * 8, hipMemcpyDeviceToDevice, ctx->stream));
            break;
    }
    checkLaunch();
    sync(ctx);
    MRMD_API_END
}

int mrmd_molecules_download(const mrmd_molecules* mc, int field, void* host, int64_t off, int64_t n)
{
    auto* m = const_cast<mrmd_molecules*>(mc);
    MRMD_API_BEGIN(m ? m->ctx : nullptr)
    REQUIRE(m && host && off >= 0 && n >= 0 && off + n <= m->size && field >= 0 && field <= MRMD_MOL_NUMATOMS,
            "invalid arguments");
    if (n == 0) return MRMD_OK;
    mrmd_ctx* ctx = m->ctx;
    const bool vec = field == MRMD_MOL_POS || field == MRMD_MOL_FORCE || field == MRMD_MOL_GRADLAMBDA;
    m->stage.ensure(ctx, (vec ? 3 : 1) * n);
    const dim3 g(gridFor(n)), b(BLOCK);
    const void* src = nullptr;
    switch (field)
    {
        case MRMD_MOL_POS:
            hipLaunchKernelGGL(k_getVec4, g, b, 0, ctx->stream, m->pos.p, m->stage.p, off, n);
            break;
        case MRMD_MOL_FORCE:
            hipLaunchKernelGGL(k_getSoA3, g, b, 0, ctx->stream, m->force.p, m->cap, m->stage.p, off, n);
            break;
        case MRMD_MOL_GRADLAMBDA:
            hipLaunchKernelGGL(k_getSoA3, g, b, 0, ctx->stream, m->grad.p, m->cap, m->stage.p, off, n);
            break;
        case MRMD_MOL_LAMBDA:
            src = m->lambda.p + off;
            break;
        case MRMD_MOL_MODLAMBDA:
            src = m->modLambda.p + off;
            break;
        case MRMD_MOL_OFFSET:
            src = m->offset.p + off;
            break;
        case MRMD_MOL_NUMATOMS:
            src = m->numAtoms.p + off;
            break;
    }
    checkLaunch();
    if (src)
        HIPCHECK(hipMemcpyAsync(host, src, size_t(n) * 8, hipMemcpyDeviceToHost, ctx->stream));
    else
        HIPCHECK(hipMemcpyAsync(host, m->stage.p, size_t(3 * n) * 8, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    MRMD_API_END
}

int mrmd_molecules_fill(mrmd_molecules* m, int field, double value)
{
    MRMD_API_BEGIN(m ? m->ctx : nullptr)
    REQUIRE(m && field >= 0 && field <= MRMD_MOL_NUMATOMS, "invalid arguments");
    mrmd_ctx* ctx = m->ctx;
    const int64_t n = m->size;
    if (n == 0) return MRMD_OK;
    if (field == MRMD_MOL_OFFSET || field == MRMD_MOL_NUMATOMS) ++m->layoutVersion;
    const dim3 g(gridFor(n)), b(BLOCK);
    switch (field)
    {
        case MRMD_MOL_POS:
            hipLaunchKernelGGL(k_fillPos, g, b, 0, ctx->stream, m->pos.p, n, value);
            break;
        case MRMD_MOL_FORCE:
        case MRMD_MOL_GRADLAMBDA:
        {
            double* p = field == MRMD_MOL_FORCE ? m->force.p : m->grad.p;
            hipLaunchKernelGGL(k_fill3, dim3(gridFor(n), 3), b, 0, ctx->stream, p, m->cap, n, value);
            break;
        }
        case MRMD_MOL_LAMBDA:
            hipLaunchKernelGGL(k_fill, g, b, 0, ctx->stream, m->lambda.p, n, value);
            break;
        case MRMD_MOL_MODLAMBDA:
            hipLaunchKernelGGL(k_fill, g, b, 0, ctx->stream, m->modLambda.p, n, value);
            break;
        case MRMD_MOL_OFFSET:
            hipLaunchKernelGGL(k_fillI64, g, b, 0, ctx->stream, m->offset.p, n, int64_t(value));
            break;
        case MRMD_MOL_NUMATOMS:
            hipLaunchKernelGGL(k_fillI64, g, b, 0, ctx->stream, m->numAtoms.p, n, int64_t(value));
            break;
    }
    checkLaunch();
    MRMD_API_END
}

}  // extern "C"

namespace
{
__global__ void k_molForEachAtom(MolView m, int64_t size)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size) return;
    m.offset[i] = i;
    m.numAtoms[i] = 1;
}
}  // namespace

extern "C" int mrmd_molecules_for_each_atom(mrmd_molecules* m, const mrmd_atoms* a)
{
    MRMD_API_BEGIN(m ? m->ctx : nullptr)
    REQUIRE(m && a, "invalid arguments");
    const int64_t size = a->nlocal + a->nghost;
    // data::Molecules molecules(2 * size) (MoleculesFromAtoms.cpp:22)
    m->size = 0;
    m->setSize(2 * size);
    if (size > 0)
    {
        hipLaunchKernelGGL(k_molForEachAtom, dim3(gridFor(size)), dim3(BLOCK), 0, m->ctx->stream, m->view(), size);
        checkLaunch();
    }
    m->nlocal = a->nlocal;
    m->nghost = a->nghost;
    ++m->layoutVersion;
    MRMD_API_END
}

int mrmd_hip::fillForceRange(mrmd_atoms* a, int64_t lo, int64_t hi)
{
    MRMD_API_BEGIN(a ? a->ctx : nullptr)
    REQUIRE(a && lo >= 0 && lo <= hi && hi <= a->size, "invalid record range");
    if (hi > lo)
        hipLaunchKernelGGL(k_fill3, dim3(gridFor(hi - lo), 3), dim3(BLOCK), 0, a->ctx->stream, a->force.p + lo, a->cap,
                           hi - lo, 0.0);
    checkLaunch();
    MRMD_API_END
}
