// rx_common.h — host-side helpers shared by the library's translation units:
// status/error plumbing for the C ABI and launch-shape utilities.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "roxy_hip.h"

namespace rx {

// Records a formatted message as this thread's last error; returns `code`.
int set_error(int code, const char* fmt, ...);

// The library's compute units (queried once per device and cached).
int device_cu_count();

}  // namespace rx

#define RX_TRY_HIP(expr)                                                                   \
    do {                                                                                   \
        hipError_t rx_e_ = (expr);                                                         \
        if (rx_e_ != hipSuccess)                                                           \
            return rx::set_error(RX_ERR_HIP, "%s failed: %s (%s:%d)", #expr,               \
                                 hipGetErrorString(rx_e_), __FILE__, __LINE__);            \
    } while (0)

#define RX_CHECK_LAUNCH()                                                                  \
    do {                                                                                   \
        hipError_t rx_e_ = hipGetLastError();                                              \
        if (rx_e_ != hipSuccess)                                                           \
            return rx::set_error(RX_ERR_HIP, "kernel launch failed: %s (%s:%d)",           \
                                 hipGetErrorString(rx_e_), __FILE__, __LINE__);            \
    } while (0)

#define RX_REQUIRE(cond, code, ...)                                                        \
    do {                                                                                   \
        if (!(cond)) return rx::set_error((code), __VA_ARGS__);                           \
    } while (0)

static inline hipStream_t rx_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

static inline size_t rx_align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Every scene array is 16-byte aligned (records are read as 16-byte words).
static inline bool rx_scene_aligned(const rx_scene& sc) {
    const void* ptrs[] = {sc.nodes, sc.header, sc.tris, sc.tri_shade, sc.materials, sc.prims, sc.prim_materials,
                          sc.lights, sc.textures};
    for (const void* p : ptrs)
        if (reinterpret_cast<uintptr_t>(p) & 15) return false;
    return true;
}

// Bump allocator over a caller-provided workspace (256-byte aligned slices).
struct rx_arena {
    char* base;
    size_t cap;
    size_t off;
    template <class T>
    T* take(size_t count) {
        off = rx_align_up(off, 256);
        T* p = reinterpret_cast<T*>(base ? base + off : nullptr);
        off += count * sizeof(T);
        return p;
    }
    void* take_bytes(size_t bytes) {
        off = rx_align_up(off, 256);
        void* p = base ? base + off : nullptr;
        off += bytes;
        return p;
    }
};
