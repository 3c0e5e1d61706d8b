// ads_common.h -- internal helpers shared by the HIP translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/adseismic.h"

namespace ads {

// ---- error state --------------------------------------------------------------------------
void set_error(const char *fmt, ...);
void clear_error();

struct Status {
    int code = ADS_OK;
};

#define ADS_FAIL(code_, ...)                                                                       \
    do {                                                                                           \
        ::ads::set_error(__VA_ARGS__);                                                             \
        return (code_);                                                                            \
    } while (0)

#define ADS_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            ADS_FAIL(ADS_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                     __LINE__);                                                                    \
    } while (0)

#define ADS_LAUNCH_CHECK()                                                                         \
    do {                                                                                           \
        hipError_t e_ = hipGetLastError();                                                         \
        if (e_ != hipSuccess)                                                                      \
            ADS_FAIL(ADS_ERR_HIP, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_),      \
                     __FILE__, __LINE__);                                                          \
    } while (0)

#define ADS_TRY(expr)                                                                              \
    do {                                                                                           \
        int rc_ = (expr);                                                                          \
        if (rc_ != ADS_OK) return rc_;                                                             \
    } while (0)

inline hipStream_t S(ads_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// Device buffer owned by a plan (setup-time allocation only).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int alloc(size_t n)
    {
        release();
        if (n == 0) return ADS_OK;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            ADS_FAIL(ADS_ERR_HIP, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
        }
        bytes = n;
        return ADS_OK;
    }
    template <class V> int upload(const std::vector<V> &v)
    {
        ADS_TRY(alloc(v.size() * sizeof(V)));
        if (!v.empty()) {
            hipError_t e = hipMemcpy(p, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice);
            if (e != hipSuccess)
                ADS_FAIL(ADS_ERR_HIP, "hipMemcpy H2D failed: %s", hipGetErrorString(e));
        }
        return ADS_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class V> V *as() const { return reinterpret_cast<V *>(p); }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

// Bump allocator over a caller-provided workspace.
struct Carve {
    char *base;
    size_t cap, off = 0;
    Carve(void *b, size_t c) : base(static_cast<char *>(b)), cap(c) {}
    void *take(size_t bytes)
    {
        size_t a = (off + 255) & ~size_t(255);
        off = a + bytes;
        return base ? base + a : nullptr;
    }
    bool ok() const { return off <= cap; }
};

// Unique-cell tables for deterministic injection: per shot, the distinct flat cells and, for
// each, the list of entries (source or receiver numbers within the shot) in original order.
struct CellLists {
    std::vector<int64_t> shot_ptr;  // [nshots+1] into cells
    std::vector<int64_t> cell;      // flat 0-based cell
    std::vector<int64_t> ent_ptr;   // [ncells+1] into ent
    std::vector<int64_t> ent;       // entry number within the shot
};

// Benchmark timing (see ads_profile_* in the header).  prof_start returns -1 when disabled.
int prof_start(int cls, hipStream_t st);
bool prof_enabled();
// alg_bytes: algorithmic bytes of the timed launches; levels: time levels they advance (each
// counts as one launch of the class)
void prof_stop(int rec, hipStream_t st, double alg_bytes, int levels = 1);

CellLists build_cell_lists(int64_t nshots, const int64_t *ptr, const std::vector<int64_t> &flat);

}  // namespace ads
