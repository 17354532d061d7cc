// Shared helpers of the gfx950 ERT library: status handling, launch helpers and
// wavefront reductions.  Wavefronts are 64 lanes wide on CDNA4.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/adt_hip.h"

namespace adt {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // four wavefronts per workgroup
constexpr int kXcds = 8;     // MI355X: 8 XCDs, workgroups are dealt round-robin across them

int set_error(int code, const char* fmt, ...);

#define ADT_HIP_CHECK(expr)                                                                    \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return ::adt::set_error(ADT_ERR_HIP, "%s failed: %s (%s:%d)", #expr,                \
                                    hipGetErrorString(_e), __FILE__, __LINE__);                \
    } while (0)

#define ADT_REQUIRE(cond, ...)                                                                 \
    do {                                                                                       \
        if (!(cond)) return ::adt::set_error(ADT_ERR_INVALID, __VA_ARGS__);                    \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Global-memory view of a device pointer.  Pointers that kernels read out of argument blocks
// in device memory are generic, and accesses through generic pointers compile to flat_*
// instructions, which count against both the vector-memory and the LDS/scalar counters: every
// wait for an LDS or scalar result then also waits for all loads in flight.  Accesses through
// the view compile to global_* instructions.  Only for pointers into global memory.  (The host
// compilation pass, which only parses kernel bodies, sees the identity.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gview(T* p) {
    return (__attribute__((address_space(1))) T*)(p);
}
#else
template <typename T>
__device__ __forceinline__ T* gview(T* p) {
    return p;
}
#endif

// s_waitcnt vmcnt(0): every vector memory operation of the wavefront has completed.  The
// compiler tracks loads whose only uses sit in optional or predicated code (halo batches past
// the first, stores into LDS for rows in range) as possibly outstanding past that code, and
// then, wherever their registers are reused, waits for all memory operations issued before
// the reuse -- the stores of earlier rows included.  Called where nothing useful is left in
// flight anyway (before the barrier of a staged product), it clears that state.
__device__ __forceinline__ void wait_vector_memory() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Workgroup b runs on XCD b % 8.  Remap so that each XCD receives chunks of kXcdChunk
// consecutive tiles (neighbouring row ranges of one system share its L2), dealt round-robin:
// XCD x takes chunks x, x + 8, x + 16, ...  Whole-range blocks per XCD would put each system on
// one or two XCDs, and late in a solve, when only the slowest systems are still iterating, most
// XCDs would idle.  The tail past the last full round of chunks is split into contiguous ranges.
// Bijective for any count.
constexpr int kXcdChunk = 16;

__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int full = n / (kXcds * kXcdChunk) * (kXcds * kXcdChunk);
    const int x = b % kXcds, i = b / kXcds;
    if (b < full) return ((i / kXcdChunk) * kXcds + x) * kXcdChunk + (i % kXcdChunk);
    const int m = n - full, bt = b - full;
    const int q = m / kXcds, rmd = m % kXcds;
    const int xt = bt % kXcds, it = bt / kXcds;
    return full + xt * q + (xt < rmd ? xt : rmd) + it;
}

// Butterfly sum over the lanes of a wavefront that share (lane % width): lanes whose ids
// differ in bits >= log2(width).  Fixed order -> bitwise repeatable.
template <int W>
__device__ __forceinline__ double wave_sum_strided(double v) {
#pragma unroll
    for (int off = W; off < kWave; off <<= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Full-wave sum (every lane receives the total), fixed butterfly order.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

inline int blocks_for(int64_t n, int per) { return static_cast<int>((n + per - 1) / per); }

}  // namespace adt
