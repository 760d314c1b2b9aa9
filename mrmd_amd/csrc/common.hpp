// Shared device/host infrastructure of the gfx950 library: context, errors, device
// buffers, launch helpers, per-kernel event timing and wave-level primitives.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mrmd_hip.h"

namespace mrmd_hip
{
constexpr int WAVE = 64;    // CDNA wavefront width
constexpr int BLOCK = 256;  // 4 waves per workgroup

// error carried through the C ABI as (code, message)
struct Error : std::runtime_error
{
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

#define HIPCHECK(call)                                                                    \
    do                                                                                    \
    {                                                                                     \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            ::mrmd_hip::fail(e_ == hipErrorOutOfMemory ? MRMD_ERR_NOMEM : MRMD_ERR_HIP,   \
                             "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                             __FILE__, __LINE__);                                         \
    } while (0)

#define REQUIRE(cond, ...)                                      \
    do                                                          \
    {                                                           \
        if (!(cond)) ::mrmd_hip::fail(MRMD_ERR_ARG, __VA_ARGS__); \
    } while (0)

struct KStat
{
    int64_t launches = 0;
    double ms = 0.0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
};
}  // namespace mrmd_hip

struct mrmd_ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    bool ownStream = false;
    std::string err;
    bool profiling = false;
    std::string profileOnly;  ///< non-empty: only timers of this name record
    std::map<std::string, mrmd_hip::KStat> stats;
    std::vector<hipEvent_t> eventPool;
    double* pinned = nullptr;  // small pinned host staging area (64 words, coherent)
    uint32_t hostTag = 0;      // tag of the last call whose result kernels post to pinned
    int64_t bytes = 0;         // device bytes held by objects of this context
    void* scratch[4] = {nullptr, nullptr, nullptr, nullptr};  // per-purpose device scratch
    size_t scratchBytes[4] = {0, 0, 0, 0};
    int refs = 1;  // the user's handle plus one per live object created on this context
};

namespace mrmd_hip
{
void ctxRetain(mrmd_ctx* c);
void ctxRelease(mrmd_ctx* c);

// owning pointer of an object to its context: the context is freed only after the last
// object created on it (mrmd_ctx_destroy drops the user's reference)
struct CtxPtr
{
    mrmd_ctx* p = nullptr;
    CtxPtr() = default;
    CtxPtr(std::nullptr_t) {}
    CtxPtr(mrmd_ctx* c) : p(c)
    {
        if (p) ctxRetain(p);
    }
    CtxPtr(const CtxPtr& o) : p(o.p)
    {
        if (p) ctxRetain(p);
    }
    CtxPtr& operator=(const CtxPtr& o) { return *this = o.p; }
    CtxPtr& operator=(mrmd_ctx* c)
    {
        if (c) ctxRetain(c);
        if (p) ctxRelease(p);
        p = c;
        return *this;
    }
    ~CtxPtr()
    {
        if (p) ctxRelease(p);
    }
    operator mrmd_ctx*() const { return p; }
    mrmd_ctx* operator->() const { return p; }
};
}  // namespace mrmd_hip

namespace mrmd_hip
{
// ---------------------------------------------------------------------------------------
// device buffer with explicit growth; contents are preserved on grow when asked to
template <class T>
struct DBuf
{
    T* p = nullptr;
    int64_t n = 0;
    mrmd_ctx* ctx = nullptr;

    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    /// capacity for at least `need` elements, growing by a quarter beyond it so that a
    /// slowly varying size does not reallocate on every call
    void ensure(mrmd_ctx* c, int64_t need, int64_t keep = 0)
    {
        if (need > n) reserve(c, need + need / 4 + 256, keep);
    }
    void swap(DBuf& o)
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
        std::swap(ctx, o.ctx);
    }

    void release()
    {
        if (p)
        {
            (void)hipFree(p);
            if (ctx) ctx->bytes -= n * int64_t(sizeof(T));
        }
        p = nullptr;
        n = 0;
    }
    // ensure capacity >= m elements; keep the first `keep` elements
    void reserve(mrmd_ctx* c, int64_t m, int64_t keep = 0)
    {
        if (m <= n) return;
        ctx = c;
        T* q = nullptr;
        HIPCHECK(hipMalloc(&q, size_t(m) * sizeof(T)));
        if (p && keep > 0)
            HIPCHECK(hipMemcpyAsync(q, p, size_t(std::min(keep, n)) * sizeof(T), hipMemcpyDeviceToDevice,
                                    c->stream));
        if (p)
        {
            HIPCHECK(hipStreamSynchronize(c->stream));
            (void)hipFree(p);
            c->bytes -= n * int64_t(sizeof(T));
        }
        p = q;
        n = m;
        c->bytes += m * int64_t(sizeof(T));
    }
};

// per-context device scratch slot (0: scan partial sums, 1: reduction partials,
// 2: small device scalars); grows, never shrinks, freed with the context
inline void* ctxScratch(mrmd_ctx* ctx, int slot, size_t bytes)
{
    if (bytes > ctx->scratchBytes[slot])
    {
        HIPCHECK(hipStreamSynchronize(ctx->stream));
        if (ctx->scratch[slot]) (void)hipFree(ctx->scratch[slot]);
        ctx->scratch[slot] = nullptr;
        size_t nb = bytes + bytes / 2 + 4096;
        HIPCHECK(hipMalloc(&ctx->scratch[slot], nb));
        ctx->scratchBytes[slot] = nb;
    }
    return ctx->scratch[slot];
}

inline int64_t growCap(int64_t need) { return need + need / 4 + 256; }

inline unsigned gridFor(int64_t n, int block = BLOCK)
{
    int64_t g = (n + block - 1) / block;
    return unsigned(g < 1 ? 1 : g);
}

// ---------------------------------------------------------------------------------------
// per-kernel event timing on the context stream
struct KTimer
{
    mrmd_ctx* ctx;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(mrmd_ctx* c, const char* n) : ctx(c), name(n)
    {
        if (!ctx->profiling || (!ctx->profileOnly.empty() && ctx->profileOnly != name)) return;
        a = take();
        b = take();
        HIPCHECK(hipEventRecord(a, ctx->stream));
    }
    ~KTimer()
    {
        if (!ctx->profiling || !a) return;
        (void)hipEventRecord(b, ctx->stream);
        ctx->stats[name].pending.push_back({a, b});
    }
    hipEvent_t take()
    {
        hipEvent_t e;
        if (!ctx->eventPool.empty())
        {
            e = ctx->eventPool.back();
            ctx->eventPool.pop_back();
        }
        else
            HIPCHECK(hipEventCreate(&e));
        return e;
    }
};

inline void flushStats(mrmd_ctx* ctx)
{
    HIPCHECK(hipStreamSynchronize(ctx->stream));
    for (auto& kv : ctx->stats)
    {
        for (auto& pr : kv.second.pending)
        {
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, pr.first, pr.second));
            kv.second.ms += ms;
            kv.second.launches += 1;
            ctx->eventPool.push_back(pr.first);
            ctx->eventPool.push_back(pr.second);
        }
        kv.second.pending.clear();
    }
}

inline void sync(mrmd_ctx* ctx) { HIPCHECK(hipStreamSynchronize(ctx->stream)); }

// Kernels post small results to the context's coherent pinned buffer as tagged 64-bit words:
// the high half is the call's tag (nextTag: a per-context counter, never 0), the low half a
// 32-bit value, each word written by one atomic 64-bit store.  The host polls until every
// word of the call carries the call's tag (waitTags), so a value is taken only from a word
// that this call wrote, whatever order the words reach host memory in; the stream is
// synchronised only as a fallback after 20 ms.  Words (uint64 at ctx->pinned, 64 of them):
// 0..3 untagged host copies (energies, after a synchronisation), 4..5 the displacement
// maximum (low, high half of the double), 8..23 the integers of readbackI32.
constexpr int DISP_WORD = 4;
constexpr int READBACK_WORD = 8;
constexpr int READBACK_INTS = 16;
constexpr int PINNED_WORDS = 64;
uint32_t nextTag(mrmd_ctx* ctx);
void waitTags(mrmd_ctx* ctx, int word, int n, uint32_t tag);
// the low halves of n tagged words (after waitTags)
inline uint32_t tagValue(const mrmd_ctx* ctx, int word)
{
    return uint32_t(__atomic_load_n(reinterpret_cast<const uint64_t*>(ctx->pinned) + word, __ATOMIC_ACQUIRE));
}
// n (<= READBACK_INTS) int32 device values to the host through the pinned buffer
void readbackI32(mrmd_ctx* ctx, const int32_t* dev, int n, int32_t* out);

__host__ __device__ inline uint64_t tagWord(uint32_t tag, uint32_t v) { return (uint64_t(tag) << 32) | v; }

inline void checkLaunch() { HIPCHECK(hipGetLastError()); }

// ---------------------------------------------------------------------------------------
// predicates (descriptor form; semantics in mrmd_hip.h)
__host__ __device__ inline bool evalPred1(const mrmd_pred1& p, double x, double y, double z)
{
    const double v = p.axis == 0 ? x : (p.axis == 1 ? y : z);
    switch (p.kind)
    {
        case MRMD_PRED_ALWAYS:
            return true;
        case MRMD_PRED_NEVER:
            return false;
        case MRMD_PRED_OPEN_RANGE:
            return v > p.a && v < p.b;
        case MRMD_PRED_SYMMETRIC_SLAB:
        {
            double absDx = fabs(v - p.a);
            return absDx >= p.b - p.tol && absDx <= p.c + p.tol;
        }
        case MRMD_PRED_SLAB_AT:
            return fabs(v - p.a) < p.b;
        case MRMD_PRED_SLAB_HY:
        {
            double absDx = fabs(v - p.a);
            if (absDx < p.b) return false;
            return absDx < p.b + p.c;
        }
        case MRMD_PRED_SLAB_CG:
            return !(fabs(v - p.a) < p.b + p.c);
        case MRMD_PRED_GREATER:
            return v > p.a;
        case MRMD_PRED_LESS:
            return v < p.a;
        default:
            return false;
    }
}

__host__ __device__ inline bool evalPred2(const mrmd_pred2& p,
                                          double xi, double yi, double zi,
                                          double xj, double yj, double zj)
{
    switch (p.op)
    {
        case MRMD_PAIR_ALWAYS:
            return true;
        case MRMD_PAIR_AND:
            return evalPred1(p.first, xi, yi, zi) && evalPred1(p.second, xj, yj, zj);
        case MRMD_PAIR_OR:
            return evalPred1(p.first, xi, yi, zi) || evalPred1(p.second, xj, yj, zj);
        default:
            return false;
    }
}

inline bool predIsSymmetric(const mrmd_pred2* p)
{
    if (!p || p->op == MRMD_PAIR_ALWAYS) return true;
    const auto& a = p->first;
    const auto& b = p->second;
    return a.kind == b.kind && a.axis == b.axis && a.a == b.a && a.b == b.b && a.c == b.c && a.tol == b.tol;
}

inline void validatePred1(const mrmd_pred1* p)
{
    if (!p) return;
    REQUIRE(p->kind >= 0 && p->kind <= MRMD_PRED_LESS, "unknown predicate kind %d", p->kind);
    REQUIRE(p->axis >= 0 && p->axis <= 2, "predicate axis %d out of range", p->axis);
}
inline void validatePred2(const mrmd_pred2* p)
{
    if (!p) return;
    REQUIRE(p->op >= 0 && p->op <= MRMD_PAIR_OR, "unknown pair predicate op %d", p->op);
    validatePred1(&p->first);
    validatePred1(&p->second);
}

// ---------------------------------------------------------------------------------------
// wave / block reductions (wave64 shuffles, then LDS across the 4 waves of a block)
__device__ inline double waveSum(double v)
{
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}
__device__ inline double waveMax(double v)
{
    for (int off = WAVE / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, WAVE));
    return v;
}
// result valid in thread 0; lds holds blockDim.x / WAVE doubles (blockDim.x a multiple of WAVE, at most BLOCK)
__device__ inline double blockSum(double v, double* lds)
{
    v = waveSum(v);
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < int(blockDim.x) / WAVE; ++i) r += lds[i];
    __syncthreads();
    return r;
}
__device__ inline double blockMax(double v, double* lds)
{
    v = waveMax(v);
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < int(blockDim.x) / WAVE; ++i) r = fmax(r, lds[i]);
    __syncthreads();
    return r;
}

template <class T>
__device__ inline T waveInclusiveScan(T v)
{
    const int lane = threadIdx.x % WAVE;
    for (int off = 1; off < WAVE; off <<= 1)
    {
        T o = __shfl_up(v, off, WAVE);
        if (lane >= off) v += o;
    }
    return v;
}

// exclusive scan of one value per thread over a BLOCK-thread workgroup; *total = block sum
template <class T>
__device__ inline T blockExclusiveScan(T v, T* lds /*[BLOCK/WAVE]*/, T* total)
{
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    T inc = waveInclusiveScan(v);
    if (lane == WAVE - 1) lds[w] = inc;
    __syncthreads();
    T waveOff = 0, tot = 0;
    for (int i = 0; i < BLOCK / WAVE; ++i)
    {
        if (i < w) waveOff += lds[i];
        tot += lds[i];
    }
    __syncthreads();
    *total = tot;
    return waveOff + inc - v;
}

// rank of this lane among the lanes of its wave whose flag is set (wave ballot + popcount)
__device__ inline int waveRank(unsigned long long mask)
{
    const unsigned lane = threadIdx.x % WAVE;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    return __popcll(mask & below);
}

// non-negative doubles order like their bit patterns
__device__ inline void atomicMaxNonNeg(double* addr, double v)
{
    atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

__device__ inline void atomicAddD(double* addr, double v) { unsafeAtomicAdd(addr, v); }

// Consecutive lanes of a wave with the same key (a cell, a histogram bin) form a run
// (records in cell order give a few runs per wave); the run's first lane does the run's
// atomic, so one atomic per run replaces one per record.  The valid lanes must be a prefix
// of the wave.
struct Run
{
    int head;  ///< lane of the run's first record
    int len;   ///< records in the run (valid on the head lane)
};

__device__ inline Run laneRun(int c, bool valid)
{
    const int lane = int(threadIdx.x) % WAVE;
    const int prev = __shfl_up(c, 1, WAVE);
    const bool head = valid && (lane == 0 || prev != c);
    const unsigned long long heads = __ballot(head);
    const unsigned long long valids = __ballot(valid);
    const unsigned long long upTo = lane == WAVE - 1 ? ~0ull : ((2ull << lane) - 1);
    Run r;
    r.head = heads & upTo ? 63 - __clzll(heads & upTo) : lane;
    const unsigned long long above = heads & ~upTo;
    const int end = above ? __builtin_ctzll(above) : __popcll(valids);
    r.len = end - lane;
    return r;
}

// XCD-aware block remap: gfx950 dispatches consecutive workgroups round-robin over 8 XCDs,
// each with its own L2.  Remapping gives every XCD a contiguous range of workgroups, so
// spatially neighbouring atoms (cell-sorted) share one L2.
__device__ inline unsigned xcdBlock()
{
    const unsigned n = gridDim.x, b = blockIdx.x;
    const unsigned x = b & 7u, k = b >> 3;
    const unsigned per = n >> 3;        // full rounds
    const unsigned rem = n & 7u;        // XCDs with one extra block
    // XCD x owns per + (x < rem) blocks, starting after the blocks of XCDs < x
    const unsigned start = x * per + (x < rem ? x : rem);
    return start + k;
}

// exclusive scan of int64 counts on the device (small helper used by binning/compaction)
void scanExclusiveI64(mrmd_ctx* ctx, const int32_t* in, int64_t* out, int64_t n, int64_t* totalDev);
void scanExclusiveI32(mrmd_ctx* ctx, const int32_t* in, int32_t* out, int64_t n, int32_t* totalDev);
// `count` arrays of n values each, consecutive in memory (array k at in + k * n), scanned
// independently; totals to totalsDev[0, count)
void scanExclusiveI32Multi(mrmd_ctx* ctx, const int32_t* in, int32_t* out, int64_t n, int count, int32_t* totalsDev);

}  // namespace mrmd_hip

// C-ABI wrapper: converts exceptions into status codes and records the message
#define MRMD_API_BEGIN(ctxexpr) \
    mrmd_ctx* api_ctx_ = (ctxexpr); \
    try                             \
    {
#define MRMD_API_END                                   \
    }                                                  \
    catch (const ::mrmd_hip::Error& e)                 \
    {                                                  \
        if (api_ctx_) api_ctx_->err = e.what();        \
        return e.code;                                 \
    }                                                  \
    catch (const std::exception& e)                    \
    {                                                  \
        if (api_ctx_) api_ctx_->err = e.what();        \
        return MRMD_ERR_HIP;                           \
    }                                                  \
    return MRMD_OK;
