// SIMT emulation runtime for the CPU build of libmatvis_hip's kernels.
//
// Each workgroup runs as one fiber per thread (ucontext).  __syncthreads()
// suspends a thread until every live thread of the block has arrived; wave
// collectives (ballot, shuffle, MFMA) suspend until every live lane of the
// 64-lane wave has deposited its operands, then each lane computes its result
// from all lanes' operands.  A block whose live threads wait at barriers that
// cannot all be released (a barrier under divergent control flow, which
// would hang a GPU) aborts with a message.  AddressSanitizer is told about
// every fiber switch.
#include <sanitizer/common_interface_defs.h>
#include <ucontext.h>

#include <cstdlib>
#include <functional>
#include <vector>

#include "hip/hip_runtime.h"

emu_idx_t threadIdx, blockIdx, blockDim, gridDim;

namespace {

constexpr size_t kStack = 256 * 1024;

enum Wait { kRun = 0, kBlock = 1, kWave = 2, kDone = 3 };

struct Fiber {
  ucontext_t ctx;
  std::vector<char> stack;
  emu_idx_t idx;
  int linear = 0;
  int wait = kRun;
  void *fake_stack = nullptr;
};

struct Block {
  std::vector<Fiber> fibers;
  ucontext_t main_ctx;
  void *main_fake = nullptr;
  int cur = -1;
  const std::function<void()> *body = nullptr;
  // wave operand slots: [wave][lane] -> up to 256 bytes
  std::vector<unsigned char> slots;
  std::vector<int> slot_gen;
};

Block g_block;

void switch_to_main() {
  Fiber &f = g_block.fibers[g_block.cur];
  __sanitizer_start_switch_fiber(&f.fake_stack, nullptr, 0);
  swapcontext(&f.ctx, &g_block.main_ctx);
  __sanitizer_finish_switch_fiber(f.fake_stack, nullptr, nullptr);
}

void trampoline() {
  __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
  (*g_block.body)();
  Fiber &f = g_block.fibers[g_block.cur];
  f.wait = kDone;
  __sanitizer_start_switch_fiber(nullptr, nullptr, 0);
  setcontext(&g_block.main_ctx);
}

int nthreads() { return (int)g_block.fibers.size(); }

}  // namespace

int emu_lane() { return g_block.fibers[g_block.cur].linear & 63; }
int emu_wave() { return g_block.fibers[g_block.cur].linear >> 6; }

void emu_syncthreads() {
  g_block.fibers[g_block.cur].wait = kBlock;
  switch_to_main();
}

void *emu_wave_slot(int lane) {
  const int w = emu_wave();
  return &g_block.slots[((size_t)w * 64 + lane) * 256];
}

bool emu_lane_live(int lane) {
  const int t = emu_wave() * 64 + lane;
  return t < nthreads() && g_block.fibers[t].wait != kDone;
}

void emu_wave_barrier() {
  g_block.fibers[g_block.cur].wait = kWave;
  switch_to_main();
}

void emu_run_block(const std::function<void()> &body) {
  const int n = blockDim.x * blockDim.y * blockDim.z;
  g_block.fibers.resize(n);
  g_block.body = &body;
  g_block.slots.assign((size_t)((n + 63) / 64) * 64 * 256, 0);
  for (int t = 0; t < n; ++t) {
    Fiber &f = g_block.fibers[t];
    f.linear = t;
    f.idx = {(unsigned)(t % blockDim.x),
             (unsigned)((t / blockDim.x) % blockDim.y),
             (unsigned)(t / (blockDim.x * blockDim.y))};
    f.wait = kRun;
    f.stack.resize(kStack);
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = f.stack.data();
    f.ctx.uc_stack.ss_size = kStack;
    f.ctx.uc_link = nullptr;
    makecontext(&f.ctx, trampoline, 0);
  }
  const int nwaves = (n + 63) / 64;
  for (;;) {
    bool progressed = false;
    for (int t = 0; t < n; ++t) {
      Fiber &f = g_block.fibers[t];
      if (f.wait != kRun) continue;
      g_block.cur = t;
      threadIdx = f.idx;
      __sanitizer_start_switch_fiber(&g_block.main_fake, f.stack.data(),
                                     kStack);
      swapcontext(&g_block.main_ctx, &f.ctx);
      __sanitizer_finish_switch_fiber(g_block.main_fake, nullptr, nullptr);
      progressed = true;
    }
    if (progressed) continue;
    int live = 0, at_block = 0;
    for (auto &f : g_block.fibers) {
      if (f.wait != kDone) ++live;
      if (f.wait == kBlock) ++at_block;
    }
    if (live == 0) break;
    bool released = false;
    // release waves whose live lanes all wait at a wave collective
    for (int w = 0; w < nwaves; ++w) {
      int wl = 0, ww = 0;
      for (int l = 0; l < 64 && w * 64 + l < n; ++l) {
        int s = g_block.fibers[w * 64 + l].wait;
        if (s != kDone) ++wl;
        if (s == kWave) ++ww;
      }
      if (ww > 0 && ww == wl) {
        for (int l = 0; l < 64 && w * 64 + l < n; ++l)
          if (g_block.fibers[w * 64 + l].wait == kWave)
            g_block.fibers[w * 64 + l].wait = kRun;
        released = true;
      }
    }
    if (released) continue;
    if (at_block == live) {
      for (auto &f : g_block.fibers)
        if (f.wait == kBlock) f.wait = kRun;
      continue;
    }
    fprintf(stderr,
            "emu: barrier divergence in block (%u,%u,%u): %d live threads, "
            "%d at __syncthreads, the rest at wave collectives\n",
            blockIdx.x, blockIdx.y, blockIdx.z, live, at_block);
    abort();
  }
}
