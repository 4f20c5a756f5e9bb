// Host stand-ins for the HIP built-ins used by the LBVH construction
// kernels, so the CPU test can run each kernel thread by thread.
#pragma once

#include <algorithm>
#include <cstdint>

#define __global__
#define __shared__ static
#define __launch_bounds__(x)
#define __ATOMIC_RELAXED_SHIM 0

namespace khs {
namespace dev {
struct ShimIndex {
    unsigned x = 0, y = 0, z = 0;
};
inline thread_local ShimIndex threadIdx, blockIdx, blockDim;
inline void __syncthreads() {}
inline int __clz(uint32_t v) { return v ? __builtin_clz(v) : 32; }
inline unsigned atomicAdd(unsigned *p, unsigned v)
{
    const unsigned o = *p;
    *p += v;
    return o;
}
inline unsigned long long atomicAdd(unsigned long long *p, unsigned long long v)
{
    const unsigned long long o = *p;
    *p += v;
    return o;
}
inline unsigned long long atomicMin(unsigned long long *p, unsigned long long v)
{
    const unsigned long long o = *p;
    *p = std::min(o, v);
    return o;
}
inline unsigned long long atomicMax(unsigned long long *p, unsigned long long v)
{
    const unsigned long long o = *p;
    *p = std::max(o, v);
    return o;
}
template <class T> inline void __hip_atomic_store(T *p, T v, int, int) { *p = v; }
template <class T> inline T __hip_atomic_load(const T *p, int, int) { return *p; }
template <class T> inline T __hip_atomic_fetch_add(T *p, T v, int, int)
{
    const T o = *p;
    *p += v;
    return o;
}
}  // namespace dev
}  // namespace khs

#ifndef __HIP_MEMORY_SCOPE_AGENT
#define __HIP_MEMORY_SCOPE_AGENT 0
#endif
#ifndef __HIP_MEMORY_SCOPE_WORKGROUP
#define __HIP_MEMORY_SCOPE_WORKGROUP 0
#endif
