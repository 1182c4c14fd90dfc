// Test harness only: runs a HIP kernel of this repository on the host so that
// its indexing and synchronisation logic can be checked on a machine without a
// GPU (tests/test_kernel_emulation.py). One workgroup runs at a time. Its lanes
// are cooperative fibers on one host thread: a lane runs until it reaches a
// barrier (__syncthreads for the workgroup, __any and __shfl_xor for its
// wavefront) or, at random, an atomic; lanes resume in a seeded random order
// each pass, so lane steps interleave at barriers and atomics. Atomics are plain
// operations between switches. It is not a model of the GPU memory system and no
// product code uses it.
#pragma once
#include <ucontext.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <random>
#include <vector>

typedef int hipError_t;
enum { hipSuccess = 0 };
typedef void* hipStream_t;
typedef void* hipMemGenericAllocationHandle_t;
inline const char* hipGetErrorString(hipError_t) { return ""; }
inline hipError_t hipGetLastError() { return 0; }
inline hipError_t hipMalloc(void** p, size_t bytes) {
  *p = malloc(bytes);
  return *p ? 0 : 2;
}
inline hipError_t hipFree(void* p) {
  free(p);
  return 0;
}
struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct EmuIdx { unsigned x = 0, y = 0, z = 0; };
inline EmuIdx threadIdx, blockIdx, blockDim, gridDim;
struct ulonglong2 { unsigned long long x, y; };
inline ulonglong2 make_ulonglong2(unsigned long long a, unsigned long long b) { return ulonglong2{a, b}; }

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__ static
#define __HIP_MEMORY_SCOPE_AGENT 0

struct EmuBarrier {
  unsigned need = 0, arrived = 0;
  uint64_t gen = 0;
};
struct EmuLane {
  ucontext_t ctx;
  std::vector<char> stack;
  EmuBarrier* waiting = nullptr;
  uint64_t wait_gen = 0;
  bool done = false;
};
struct EmuBlock {
  ucontext_t sched;
  std::vector<EmuLane> lanes;
  EmuBarrier block;
  std::vector<EmuBarrier> waves;
  std::vector<int64_t> wave_val;  // 64 values per wavefront
  unsigned cur = 0;
  std::function<void()> body;
};
inline EmuBlock* emu_blk = nullptr;
inline uint64_t emu_seed = 12345;

inline void emu_wait(EmuBarrier& b) {
  const uint64_t g = b.gen;
  if (++b.arrived == b.need) {
    b.arrived = 0;
    ++b.gen;  // releases every lane waiting on this generation
    return;
  }
  EmuLane& me = emu_blk->lanes[emu_blk->cur];
  me.waiting = &b;
  me.wait_gen = g;
  swapcontext(&me.ctx, &emu_blk->sched);
}

// Before an atomic a lane gives way to the others at random (seeded), so that
// lane steps interleave at atomic boundaries as well as at barriers.
inline uint64_t emu_yield_state = 0x9E3779B97F4A7C15ull;
inline void emu_maybe_yield() {
  if (!emu_blk) return;
  emu_yield_state ^= emu_yield_state << 13;
  emu_yield_state ^= emu_yield_state >> 7;
  emu_yield_state ^= emu_yield_state << 17;
  if (emu_yield_state % 3 != 0) return;
  EmuLane& me = emu_blk->lanes[emu_blk->cur];
  me.waiting = nullptr;
  swapcontext(&me.ctx, &emu_blk->sched);
}

inline void __syncthreads() { emu_wait(emu_blk->block); }
inline int __any(int pred) {
  const unsigned w = threadIdx.x / 64, l = threadIdx.x % 64;
  int64_t* v = &emu_blk->wave_val[64 * (size_t)w];
  v[l] = pred != 0;
  emu_wait(emu_blk->waves[w]);
  int r = 0;
  for (unsigned i = 0; i < 64 && 64 * w + i < blockDim.x; ++i) r |= (int)v[i];
  emu_wait(emu_blk->waves[w]);
  return r;
}
template <class T>
T __shfl_xor(T x, int m) {
  const unsigned w = threadIdx.x / 64, l = threadIdx.x % 64;
  int64_t* v = &emu_blk->wave_val[64 * (size_t)w];
  v[l] = (int64_t)x;
  emu_wait(emu_blk->waves[w]);
  T r = (T)v[l ^ (unsigned)m];
  emu_wait(emu_blk->waves[w]);
  return r;
}
// Wavefront exchange: every lane of the wavefront publishes a value, then reads.
inline int64_t* emu_wave_publish(int64_t x) {
  const unsigned w = threadIdx.x / 64, l = threadIdx.x % 64;
  int64_t* v = &emu_blk->wave_val[64 * (size_t)w];
  v[l] = x;
  emu_wait(emu_blk->waves[w]);
  return v;
}
inline void emu_wave_done() { emu_wait(emu_blk->waves[threadIdx.x / 64]); }
inline unsigned emu_wave_lanes() {
  const unsigned w = threadIdx.x / 64;
  return 64 * w + 64 <= blockDim.x ? 64u : blockDim.x - 64 * w;
}
inline uint64_t __ballot(int pred) {
  const int64_t* v = emu_wave_publish(pred != 0);
  uint64_t r = 0;
  for (unsigned i = 0; i < emu_wave_lanes(); ++i) r |= (uint64_t)(v[i] != 0) << i;
  emu_wave_done();
  return r;
}
inline int __all(int pred) {
  const int64_t* v = emu_wave_publish(pred != 0);
  int r = 1;
  for (unsigned i = 0; i < emu_wave_lanes(); ++i) r &= (int)(v[i] != 0);
  emu_wave_done();
  return r;
}
template <class T>
T __shfl(T x, int src) {
  const int64_t* v = emu_wave_publish((int64_t)x);
  T r = (T)v[(unsigned)src % 64];
  emu_wave_done();
  return r;
}
// lane l reads lane l - delta (its own value for l < delta)
template <class T>
T __shfl_up(T x, unsigned delta) {
  const unsigned l = threadIdx.x % 64;
  const int64_t* v = emu_wave_publish((int64_t)x);
  T r = l >= delta ? (T)v[l - delta] : x;
  emu_wave_done();
  return r;
}
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
// every lane of the wavefront is active where the kernels call these
#define __builtin_amdgcn_readfirstlane(x) __shfl((x), 0)
#define __builtin_amdgcn_wave_barrier() emu_wave_done()
#define __builtin_amdgcn_fence(order, scope) ((void)0)
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
struct uint4 { uint32_t x, y, z, w; };
inline uint4 make_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return uint4{a, b, c, d}; }
inline int __clz(int x) { return x ? __builtin_clz((unsigned)x) : 32; }
template <class T> T min(T a, T b) { return a < b ? a : b; }
template <class T> T max(T a, T b) { return a < b ? b : a; }
template <class T, class U> T atomicAdd(T* p, U v) {
  emu_maybe_yield();
  T o = *p;
  *p = (T)(o + (T)v);
  return o;
}
template <class T, class U> T atomicExch(T* p, U v) {
  emu_maybe_yield();
  T o = *p;
  *p = (T)v;
  return o;
}
template <class T, class U, class V> T atomicCAS(T* p, U cmp, V val) {
  emu_maybe_yield();
  T o = *p;
  if (o == (T)cmp) *p = (T)val;
  return o;
}
template <class T, class U> T atomicMin(T* p, U v) {
  emu_maybe_yield();
  T o = *p;
  if ((T)v < o) *p = (T)v;
  return o;
}
template <class T, class U> T atomicOr(T* p, U v) {
  emu_maybe_yield();
  T o = *p;
  *p = (T)(o | (T)v);
  return o;
}
template <class T, class U> T atomicMax(T* p, U v) {
  emu_maybe_yield();
  T o = *p;
  if (o < (T)v) *p = (T)v;
  return o;
}
#define __hip_atomic_load(p, o, s) (*(p))

// The workgroup to run (others are skipped; a persistent kernel's one workgroup
// then takes every work item), or -1 for all, one after another.
inline int emu_only_block = -1;

inline void emu_lane_entry() {
  emu_blk->body();
  emu_blk->lanes[emu_blk->cur].done = true;
  swapcontext(&emu_blk->lanes[emu_blk->cur].ctx, &emu_blk->sched);
}

template <class... A, class... B>
void hipLaunchKernelGGL(void (*k)(A...), dim3 g, dim3 b, int, hipStream_t, B... args) {
  gridDim.x = g.x;
  gridDim.y = g.y;
  blockDim.x = b.x;
  std::mt19937_64 rng(emu_seed);
  emu_yield_state = emu_seed * 0x9E3779B97F4A7C15ull | 1;
  for (unsigned bj = 0; bj < g.x * g.y; ++bj) {
    const unsigned bi = bj % g.x;
    if (emu_only_block >= 0 && (int)bi != emu_only_block) continue;
    blockIdx.x = bi;
    blockIdx.y = bj / g.x;
    EmuBlock blk;
    emu_blk = &blk;
    const unsigned T = b.x, W = (T + 63) / 64;
    blk.block.need = T;
    blk.waves.resize(W);
    for (unsigned w = 0; w < W; ++w) blk.waves[w].need = (w + 1 < W || T % 64 == 0) ? 64 : T % 64;
    blk.wave_val.assign(64 * (size_t)W, 0);
    blk.body = [&]() { k(args...); };
    blk.lanes.resize(T);
    for (unsigned t = 0; t < T; ++t) {
      EmuLane& L = blk.lanes[t];
      L.stack.resize(1 << 16);
      getcontext(&L.ctx);
      L.ctx.uc_stack.ss_sp = L.stack.data();
      L.ctx.uc_stack.ss_size = L.stack.size();
      L.ctx.uc_link = nullptr;
      makecontext(&L.ctx, emu_lane_entry, 0);
    }
    std::vector<unsigned> order(T);
    for (unsigned t = 0; t < T; ++t) order[t] = t;
    for (unsigned left = T; left;) {
      std::shuffle(order.begin(), order.end(), rng);
      bool moved = false;
      for (unsigned t : order) {
        EmuLane& L = blk.lanes[t];
        if (L.done || (L.waiting && L.waiting->gen == L.wait_gen)) continue;
        L.waiting = nullptr;
        blk.cur = t;
        threadIdx.x = t;
        swapcontext(&blk.sched, &L.ctx);
        moved = true;
        if (L.done) --left;
      }
      if (!moved) {
        fprintf(stderr, "emulation: every lane waits (a barrier not reached by all lanes)\n");
        abort();
      }
    }
    emu_blk = nullptr;
  }
}
