// Lane pool of the host HIP emulation (see hip/hip_runtime.h).
#include "hip/hip_runtime.h"

thread_local EmuIdx threadIdx, blockIdx;
thread_local dim3 blockDim, gridDim;

namespace emu {
namespace {
struct Pool {
  int nlanes = 0;
  std::vector<std::thread> th;
  std::barrier<>* bar = nullptr;
  std::function<void()> body;
  dim3 grid, block;
  unsigned cur_block = 0;
  std::atomic<bool> quit{false};
  double dslot[1024];
  int islot[1024];
  unsigned long long ballot_mask = 0;
  std::mutex m;
};
Pool* g_pool = nullptr;

void lane_main(int lane) {
  Pool& P = *g_pool;
  threadIdx.x = lane;
  while (true) {
    P.bar->arrive_and_wait();  // start of a block (or shutdown)
    if (P.quit.load()) return;
    blockIdx.x = P.cur_block % P.grid.x;
    blockIdx.y = P.cur_block / P.grid.x;
    blockDim = P.block;
    gridDim = P.grid;
    if ((unsigned)lane < P.block.x) P.body();
    P.bar->arrive_and_wait();  // end of block
  }
}
}  // namespace

// Lanes beyond the block size idle at the block barriers; only the first
// block.x lanes run the body.  __syncthreads uses a second barrier sized to
// the active lanes.
static std::barrier<>* g_sync = nullptr;
static unsigned g_sync_n = 0;
// one barrier per 64-lane wavefront for cross-lane operations
static std::vector<std::barrier<>*> g_wave;
static std::barrier<>& wave_bar() { return *g_wave[threadIdx.x / 64]; }
static unsigned wave_size(unsigned l) {
  const unsigned w0 = l & ~63u;
  return std::min<unsigned>(64, g_sync_n - w0);
}

void run_grid(dim3 grid, dim3 block, const std::function<void()>& body) {
  const int lanes = 256;
  if (block.x > 256) {
    std::fprintf(stderr, "emu: block size %u > 256 not supported\n", block.x);
    std::abort();
  }
  if (!g_pool) {
    g_pool = new Pool();
    g_pool->nlanes = lanes;
    g_pool->bar = new std::barrier<>(lanes + 1);
    for (int i = 0; i < lanes; ++i) g_pool->th.emplace_back(lane_main, i);
  }
  if (g_sync_n != block.x) {
    delete g_sync;
    g_sync = new std::barrier<>(block.x);
    g_sync_n = block.x;
    for (auto* b : g_wave) delete b;
    g_wave.clear();
    for (unsigned w0 = 0; w0 < block.x; w0 += 64) g_wave.push_back(new std::barrier<>(std::min<unsigned>(64, block.x - w0)));
  }
  Pool& P = *g_pool;
  P.body = body;
  P.grid = grid;
  P.block = block;
  for (unsigned b = 0; b < grid.x * grid.y; ++b) {
    P.cur_block = b;
    P.bar->arrive_and_wait();
    P.bar->arrive_and_wait();
  }
}

void syncthreads() { g_sync->arrive_and_wait(); }

double shfl_xor(double v, int off) {
  const unsigned l = threadIdx.x;
  g_pool->dslot[l] = v;
  wave_bar().arrive_and_wait();
  const unsigned src = l ^ (unsigned)off;
  const double r = src < g_sync_n ? g_pool->dslot[src] : 0.0;
  wave_bar().arrive_and_wait();
  return r;
}

int shfl_xor(int v, int off) {
  const unsigned l = threadIdx.x;
  g_pool->islot[l] = v;
  wave_bar().arrive_and_wait();
  const unsigned src = l ^ (unsigned)off;
  const int r = src < g_sync_n ? g_pool->islot[src] : 0;
  wave_bar().arrive_and_wait();
  return r;
}

int syncthreads_or(int p) {
  const unsigned l = threadIdx.x;
  g_pool->islot[l] = p ? 1 : 0;
  g_sync->arrive_and_wait();
  int r = 0;
  for (unsigned i = 0; i < g_sync_n; ++i) r |= g_pool->islot[i];
  g_sync->arrive_and_wait();
  return r;
}

unsigned long long ballot(bool p) {
  const unsigned l = threadIdx.x;
  g_pool->islot[l] = p ? 1 : 0;
  wave_bar().arrive_and_wait();
  unsigned long long m = 0;
  const unsigned w0 = l & ~63u;
  for (unsigned i = w0; i < w0 + 64 && i < g_sync_n; ++i)
    if (g_pool->islot[i]) m |= 1ull << (i - w0);
  wave_bar().arrive_and_wait();
  return m;
}
}  // namespace emu
