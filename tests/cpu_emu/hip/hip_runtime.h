// Host stand-in for <hip/hip_runtime.h>, used only to compile libmatvis_hip's
// sources for the CPU so that every kernel's memory accesses can be checked
// under AddressSanitizer.
// each workgroup runs as fibers with real barrier semantics, and wave
// intrinsics and MFMA are wave collectives computed from the operand layouts
// the kernels assume (emu_runtime.cpp).
#pragma once

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <type_traits>

#define __global__
#define __device__
#define __host__
#define __shared__ static
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_idx_t {
  unsigned x, y, z;
};
extern emu_idx_t threadIdx, blockIdx, blockDim, gridDim;

typedef int hipError_t;
constexpr hipError_t hipSuccess = 0;
typedef struct emu_stream *hipStream_t;
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipGetDevice(int *d) {
  *d = 0;
  return hipSuccess;
}
// A file-scope __device__ array is ordinary host memory here, so its symbol
// address is the array's own address.
#define HIP_SYMBOL(sym) (reinterpret_cast<const void *>(&(sym)))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) {
  *p = const_cast<void *>(sym);
  return hipSuccess;
}
// A small emulated device: 3 CUs x 2 blocks = 6 slots, rounded by callers to
// the 8-XCD granularity, so persistent kernels loop over several items.
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 };
inline hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) {
  *v = 3;
  return hipSuccess;
}
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n,
                                                               const void *,
                                                               int, size_t) {
  *n = 2;
  return hipSuccess;
}

// SIMT runtime (emu_runtime.cpp)
void emu_syncthreads();
int emu_lane();
void *emu_wave_slot(int lane);
bool emu_lane_live(int lane);
void emu_wave_barrier();

inline void __syncthreads() { emu_syncthreads(); }
// LDS-only fences and the bare barrier: the emulated waves see memory in
// program order, so the fences are no-ops and the barrier is the block's.
#define __builtin_amdgcn_fence(order, scope, ...) ((void)0)
#define __builtin_amdgcn_s_barrier() emu_syncthreads()
inline unsigned long long __ballot(int p) {
  *(int *)emu_wave_slot(emu_lane()) = p ? 1 : 0;
  emu_wave_barrier();
  unsigned long long m = 0;
  for (int l = 0; l < 64; ++l)
    if (emu_lane_live(l) && *(int *)emu_wave_slot(l)) m |= 1ull << l;
  emu_wave_barrier();
  return m;
}
template <class T> inline T __shfl_down(T v, unsigned off, int width = 64) {
  static_assert(sizeof(T) <= 256, "slot size");
  memcpy(emu_wave_slot(emu_lane()), &v, sizeof(T));
  emu_wave_barrier();
  const int l = emu_lane();
  T r = v;
  if ((unsigned)(l % width) + off < (unsigned)width)
    memcpy(&r, emu_wave_slot(l + (int)off), sizeof(T));
  emu_wave_barrier();
  return r;
}
template <class T> inline T __shfl_xor(T v, int mask, int width = 64) {
  static_assert(sizeof(T) <= 256, "slot size");
  memcpy(emu_wave_slot(emu_lane()), &v, sizeof(T));
  emu_wave_barrier();
  const int l = emu_lane(), src = l ^ mask;
  T r = v;
  if (src < 64 && (src / width) == (l / width) && emu_lane_live(src))
    memcpy(&r, emu_wave_slot(src), sizeof(T));
  emu_wave_barrier();
  return r;
}
// v_permlane32_swap as measured on gfx950: result [0] is x of lanes 0-31 in
// lanes 0-31 and y of lanes 0-31 in lanes 32-63; result [1] is x of lanes
// 32-63 in lanes 0-31 and y of lanes 32-63 in lanes 32-63.
struct emu_u32x2 {
  unsigned v[2];
  unsigned operator[](int i) const { return v[i]; }
};
inline emu_u32x2 emu_permlane32_swap(unsigned x, unsigned y) {
  const int l = emu_lane();
  const unsigned me[2] = {x, y};
  memcpy(emu_wave_slot(l), me, sizeof me);
  emu_wave_barrier();
  unsigned other[2];
  emu_u32x2 r;
  if (l < 32) {
    memcpy(other, emu_wave_slot(l + 32), sizeof other);
    r.v[0] = x;
    r.v[1] = other[0];
  } else {
    memcpy(other, emu_wave_slot(l - 32), sizeof other);
    r.v[0] = other[1];
    r.v[1] = y;
  }
  emu_wave_barrier();
  return r;
}
#define __builtin_amdgcn_permlane32_swap(x, y, fi, bc) emu_permlane32_swap(x, y)
// v_mov_b32_dpp with a quad_perm control (ctrl < 0x100, the only DPP form
// the sources use): lane l takes the value of lane 4 (l / 4) + sel_l of its
// quad, sel_l = bits 2 (l % 4) .. 2 (l % 4) + 1 of ctrl.
inline int emu_mov_dpp(int v, int ctrl) {
  assert(ctrl >= 0 && ctrl < 0x100);
  const int l = emu_lane();
  memcpy(emu_wave_slot(l), &v, sizeof v);
  emu_wave_barrier();
  const int src = (l & ~3) | ((ctrl >> (2 * (l & 3))) & 3);
  int r = v;
  if (emu_lane_live(src)) memcpy(&r, emu_wave_slot(src), sizeof r);
  emu_wave_barrier();
  return r;
}
#define __builtin_amdgcn_mov_dpp(v, ctrl, rm, bm, bc) emu_mov_dpp(v, ctrl)
// v_perm_b32: byte i of the result is byte sel_i of the 8 bytes {s0, s1}
// (s1 in bytes 0-3, s0 in 4-7); selector 12 gives 0 and 13-255 give 0xff
// (8-11, the sign replications, are not used by the sources)
inline unsigned emu_perm(unsigned s0, unsigned s1, unsigned sel) {
  const unsigned long long both = ((unsigned long long)s0 << 32) | s1;
  unsigned r = 0;
  for (int i = 0; i < 4; ++i) {
    const unsigned b = (sel >> (8 * i)) & 0xFFu;
    assert(b < 8 || b >= 12);
    const unsigned byte =
        b < 8 ? (unsigned)(both >> (8 * b)) & 0xFFu : (b == 12 ? 0u : 0xFFu);
    r |= byte << (8 * i);
  }
  return r;
}
#define __builtin_amdgcn_perm(s0, s1, sel) emu_perm(s0, s1, sel)
// global_load_lds: lane l copies `size` bytes from its source to byte
// off + size * l of the wave's LDS destination.  Waves see memory in program
// order here, so waits on the copies are no-ops.
inline void emu_global_load_lds(const void *src, void *dst, unsigned size) {
  memcpy((char *)dst + (size_t)emu_lane() * size, src, size);
}
#define __builtin_amdgcn_global_load_lds(src, dst, size, off, aux) \
  emu_global_load_lds(src, (char *)(dst) + (off), size)
#define __builtin_amdgcn_s_waitcnt(imm) ((void)0)
#define __builtin_amdgcn_s_setprio(p) ((void)0)
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
// blocks run one at a time here; the atomic keeps the same contract anyway
inline int atomicAdd(int *p, int v) {
  return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
}
inline unsigned __float_as_uint(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  return u;
}
inline float __uint_as_float(unsigned u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
inline void sincospi(double x, double *s, double *c) {
  *s = sin(M_PI * x);
  *c = cos(M_PI * x);
}
inline void sincospif(float x, float *s, float *c) {
  *s = sinf((float)M_PI * x);
  *c = cosf((float)M_PI * x);
}
using std::max;
using std::min;

// MFMA as wave collectives, with the operand layouts of gfx950 (checked on
// the hardware with a one-wave probe).
// 16x16x4 (f64, f32): A[i][k] comes from lane i + 16 k, B[k][j] from lane
// j + 16 k; register r of lane l holds D column l % 16 and row
// 4 (l / 16) + r for f32 but (l / 16) + 4 r for f64.
template <class S, class C> inline C emu_mfma_16x16x4(S a, S b, C c) {
  struct Op {
    S a, b;
  } me{a, b};
  static_assert(sizeof(Op) <= 256, "slot size");
  const int l = emu_lane();
  memcpy(emu_wave_slot(l), &me, sizeof me);
  emu_wave_barrier();
  C d = c;
  for (int r = 0; r < 4; ++r) {
    const int i = sizeof(S) == 8 ? (l / 16) + 4 * r : 4 * (l / 16) + r;
    const int j = l % 16;
    double dot = 0.0;
    for (int k = 0; k < 4; ++k) {
      Op oa, ob;
      memcpy(&oa, emu_wave_slot(i + 16 * k), sizeof oa);
      memcpy(&ob, emu_wave_slot(j + 16 * k), sizeof ob);
      dot += (double)oa.a * (double)ob.b;
    }
    d[r] = (std::remove_reference_t<decltype(c[r])>)((double)c[r] + dot);
  }
  emu_wave_barrier();
  return d;
}
// 32x32x16 (bf16, f16; 8 per lane): A[i][k] is element k % 8 of lane
// i + 32 (k / 8), B[k][j] element k % 8 of lane j + 32 (k / 8); lane l holds
// D rows 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32.
template <class V, class C> inline C emu_mfma_32x32x16(V a, V b, C c) {
  struct Op {
    V a, b;
  } me{a, b};
  const int l = emu_lane();
  memcpy(emu_wave_slot(l), &me, sizeof me);
  emu_wave_barrier();
  // the 16 products of one instruction are summed exactly and rounded once
  // into the FP32 accumulator
  C d = c;
  for (int r = 0; r < 16; ++r) {
    const int i = 8 * (r / 4) + 4 * (l / 32) + r % 4, j = l % 32;
    double dot = 0.0;
    for (int k = 0; k < 16; ++k) {
      Op oa, ob;
      memcpy(&oa, emu_wave_slot(i + 32 * (k / 8)), sizeof oa);
      memcpy(&ob, emu_wave_slot(j + 32 * (k / 8)), sizeof ob);
      dot += (double)(float)oa.a[k % 8] * (double)(float)ob.b[k % 8];
    }
    d[r] = (float)((double)c[r] + dot);
  }
  emu_wave_barrier();
  return d;
}
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) \
  emu_mfma_16x16x4(a, b, c)
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) \
  emu_mfma_16x16x4(a, b, c)
#define __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, x, y, z) \
  emu_mfma_32x32x16(a, b, c)
#define __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, x, y, z) \
  emu_mfma_32x32x16(a, b, c)

struct emu_cfg_t {
  dim3 grid, block;
};
inline emu_cfg_t emu_cfg(dim3 g, dim3 b, size_t = 0, hipStream_t = nullptr) {
  return {g, b};
}
inline emu_cfg_t emu_cfg(dim3 g, dim3 b, size_t, void *) { return {g, b}; }

void emu_run_block(const std::function<void()> &body);

// Runs every block in turn; within a block, threads are fibers (see
// emu_runtime.cpp).
template <class... P, class... A>
void emu_launch(void (*k)(P...), emu_cfg_t cfg, A... args) {
  gridDim = {cfg.grid.x, cfg.grid.y, cfg.grid.z};
  blockDim = {cfg.block.x, cfg.block.y, cfg.block.z};
  const std::function<void()> body = [&]() { k(args...); };
  for (unsigned bz = 0; bz < cfg.grid.z; ++bz)
    for (unsigned by = 0; by < cfg.grid.y; ++by)
      for (unsigned bx = 0; bx < cfg.grid.x; ++bx) {
        blockIdx = {bx, by, bz};
        emu_run_block(body);
      }
}
