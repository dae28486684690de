// common.hpp -- shared host/device helpers of libpmx (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <algorithm>

#include "../../include/pmx.h"

namespace pmx {

constexpr int kWave = 64;
constexpr int kXcds = 8;  // MI355X: workgroups go round-robin over 8 XCDs, each with its own L2

// XCD-aware block order: workgroup b runs on XCD b % 8, so give each XCD a
// contiguous range of logical blocks (neighbouring tiles share halo data
// in that XCD's L2).  A bijection on [0, nb).
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb)
{
    const int64_t q = nb / kXcds, r = nb % kXcds, x = b % kXcds, i = b / kXcds;
    return x * q + min(x, r) + i;
}

// Thread-local last-error text behind pmx_last_error().
void set_error(const char *fmt, ...);
void clear_error();

inline int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    set_error("%s", buf);
    return code;
}

// Checks the launch just issued; returns PMX_ERR_HIP with the HIP message.
inline int check_launch(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(PMX_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return PMX_OK;
}

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Carves 256-byte aligned sub-buffers out of the caller's workspace (which
// may itself have any alignment: need() includes the slack to align it).
struct Carver {
    char *base;
    size_t cap, used = 0;
    Carver(void *b, size_t c)
        : base(b ? reinterpret_cast<char *>(align_up(reinterpret_cast<uintptr_t>(b), 256)) : nullptr), cap(c)
    {
    }
    template <class T> T *take(size_t count)
    {
        size_t off = align_up(used, 256);
        used = off + count * sizeof(T);
        return base ? reinterpret_cast<T *>(base + off) : nullptr;
    }
    size_t need() const { return align_up(used, 256) + 256; }
};

int validate_geom(const pmx_geom *g);

// Optional per-kernel timing (pmx_profile_*): when enabled, every launch
// wrapped in a ProfScope records HIP events on its stream before and after.
void prof_mark(const char *name, hipStream_t s, bool start);

struct ProfScope {
    const char *name;
    hipStream_t s;
    ProfScope(const char *n, hipStream_t st) : name(n), s(st) { prof_mark(name, s, true); }
    ~ProfScope() { prof_mark(name, s, false); }
};

}  // namespace pmx
