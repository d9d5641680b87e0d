// TEST INFRASTRUCTURE -- CPU emulation of the HIP kernel model, used only by
// tests/test_emulated_kernels.py to run the library's own .hip sources on a
// machine without a GPU (optionally under AddressSanitizer).  The product
// library never includes or links this.
//
// Every work-item of a workgroup is a fiber (ucontext) on one host thread;
// __syncthreads and each half of a __shfl_xor switch back to a scheduler
// that runs the workgroup's items round-robin, so all items meet at every
// barrier.  Workgroups of a grid run one after another, so __shared__
// (= static) storage is per workgroup.  Kernel arguments are copied at launch
// (as a real launch does), and launches made during stream capture are
// recorded and replayed by the graph launch.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
using std::isfinite;
using std::sqrt;
using std::fabs;

#undef __global__
#undef __device__
#undef __host__
#undef __shared__
#undef __forceinline__
#undef __launch_bounds__
#define __global__
#define __device__
#define __host__
#define __shared__ static
#define __forceinline__ inline
#define __launch_bounds__(...)
#ifndef HIP_SYMBOL
#define HIP_SYMBOL(X) X
#endif
#ifndef __align__
#define __align__(n) __attribute__((aligned(n)))
#endif

namespace emu {
struct Idx { unsigned x, y, z; };
extern thread_local Idx tIdx, bIdx;
extern Idx gDim, bDim;
extern thread_local int lane;
void sync_block();
double shfl(double v, int src);
void launch_fn(std::function<void()> f, unsigned grid, unsigned block);
unsigned long long clock();
template <class F, class... A>
void launch(F k, dim3 g, dim3 b, A... args)
{
  launch_fn([k, args...]() { k(args...); }, g.x, b.x);
}
}  // namespace emu

#define threadIdx emu::tIdx
#define blockIdx emu::bIdx
#define gridDim emu::gDim
#define blockDim emu::bDim

inline void __syncthreads() { emu::sync_block(); }
template <class T>
inline T __shfl_xor(T v, int off, int width = 64)
{
  const int l = emu::lane;
  const int src = (l & ~(width - 1)) | ((l & (width - 1)) ^ off);
  return (T)emu::shfl((double)v, src);
}
inline unsigned long long wall_clock64() { return emu::clock(); }

#define hipLaunchKernelGGL(K, G, B, SHM, S, ...) emu::launch(K, dim3(G), dim3(B), __VA_ARGS__)
inline int __builtin_amdgcn_readfirstlane(int v) { return v; }
template <class T> inline T __builtin_nontemporal_load(const T *p) { return *p; }
template <class T> inline void __builtin_nontemporal_store(T v, T *p) { *p = v; }
