// Shared host/device helpers for the jaxtrace_amd HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include "../../include/jaxtrace_amd.h"

namespace jt {

// Record a failure message for jt_last_error() and return `code`.
int fail(int code, const char* fmt, ...);

// Check a HIP call; on error record it and return JT_EHIP from the caller.
#define JT_HIP(call)                                                        \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess)                                               \
            return ::jt::fail(JT_EHIP, "%s failed: %s (%s:%d)", #call,      \
                              hipGetErrorString(_e), __FILE__, __LINE__);   \
    } while (0)

#define JT_REQUIRE(cond, ...)                                               \
    do {                                                                    \
        if (!(cond)) return ::jt::fail(JT_EINVAL, __VA_ARGS__);             \
    } while (0)

// Carve aligned sub-buffers out of a caller-provided workspace.
struct Arena {
    char* base;
    size_t cap, used;
    explicit Arena(void* p, size_t c) : base((char*)p), cap(c), used(0) {}
    template <class T> T* take(size_t count) {
        size_t off = (used + 255) & ~(size_t)255;
        used = off + count * sizeof(T);
        return (base && used <= cap) ? (T*)(base + off) : nullptr;
    }
    bool ok() const { return used <= cap; }
};

// ---- device helpers ----------------------------------------------------
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(((uint32_t)h) << 16);
}
// A dword whose high half is a bfloat16 and whose low half is zero is that
// value's float32 bit pattern.
__device__ __forceinline__ float bf16_hi_to_f32(uint32_t w) {
    return __uint_as_float(w);
}
__device__ __forceinline__ float f16_to_f32(uint16_t h) {
    return (float)__builtin_bit_cast(_Float16, h);
}

#ifndef JT_FMOD_OPAQUE
#define JT_FMOD_OPAQUE 1
#endif
// jnp.mod on floats: fmod, then add the divisor when the signs differ.
// For n > 0 the common cases are exact without fmod: a in [0, n) gives a;
// a in (-n, 0) gives fl(a + n) (fmod returns a itself); a in [n, 2n) gives
// a - n, which is exact (Sterbenz).  Other inputs take the general path.
__device__ __forceinline__ float jmodf(float a, float n) {
    if (n > 0.0f) {
        if (a >= 0.0f && a < n) return a;
        if (a < 0.0f && a > -n) return a + n;
        if (a >= n && a < 2.0f * n) return a - n;
    }
    // the rare general case: the opaque copy of n keeps the compiler from
    // hoisting fmodf's divisor set-up out of the callers' loops, where it
    // would hold registers for every iteration
    float m = n;
#if JT_FMOD_OPAQUE
    __asm__ volatile("" : "+v"(m));
#endif
    float r = fmodf(a, m);
    if (r != 0.0f && ((r < 0.0f) != (m < 0.0f))) r = r + m;
    return r;
}
// jnp.clip (minimum(maximum(x, lo), hi)) keeping NaN.
__host__ __device__ __forceinline__ float clipf(float x, float lo, float hi) {
    float y = x < lo ? lo : x;
    return y > hi ? hi : y;
}

}  // namespace jt
