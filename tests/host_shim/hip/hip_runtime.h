// Host stand-in for the HIP keywords the element and assembly code uses, so that
// tests/asm_host.cpp can compile that code with g++ (and its sanitizers) and run a
// kernel body on the CPU: one std::thread per GPU thread of a block, __syncthreads as a
// barrier of the block's threads.  Test infrastructure only; nothing here runs on a GPU.
#pragma once
#include <cmath>
#include <math.h>

#define __host__
#define __device__
#define __global__
#define __shared__
#define __forceinline__ inline
#define __launch_bounds__(...)

struct HostDim3 { unsigned x, y, z; };
extern thread_local HostDim3 threadIdx, blockIdx, blockDim;

void host_syncthreads();
#define __syncthreads() host_syncthreads()

inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_RELAXED); }
using std::isfinite;
