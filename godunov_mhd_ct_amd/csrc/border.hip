// border.hip -- builds the CorrectEmfAtBlockBorder gather list (EmfFix records, 8 source slots
// each) of the mesh in a context on the device, once per mesh: a count pass per workgroup of
// border slots, an exclusive scan of the counts, and an emit pass that writes each workgroup's
// records in slot order.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <string>
#include <vector>

#include "border.hpp"
#include "ctx.hpp"

namespace mhdct
{

namespace
{

constexpr int kThreads = 256;

__device__ inline bool
slot_needs_fix(const DevMesh & m, long t, long total, long & dst, long src[8], int & nsrc, int & cnt, double & ratio)
{
  if (t >= total)
    return false;
  const int per = border_slots_per_block(m.dim, m.b);
  const int o = int(t / per);
  return border_sources(m, o, int(t - long(o) * per), dst, src, nsrc, cnt, ratio);
}

__global__ void __launch_bounds__(kThreads)
k_border_count(DevMesh m, long total, int * __restrict__ counts)
{
  __shared__ int wsum[kThreads / 64];
  long           dst, src[8];
  int            nsrc, cnt;
  double         ratio;
  const long     t = long(blockIdx.x) * kThreads + threadIdx.x;
  const bool     f = slot_needs_fix(m, t, total, dst, src, nsrc, cnt, ratio);
  const int      c = __popcll(__ballot(f));
  if ((threadIdx.x & 63) == 0)
    wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0)
  {
    int s = 0;
    for (int w = 0; w < kThreads / 64; ++w)
      s += wsum[w];
    counts[blockIdx.x] = s;
  }
}

__global__ void __launch_bounds__(kThreads)
k_border_emit(DevMesh m, long total, const int * __restrict__ offsets, EmfFix * __restrict__ fix,
              int64_t * __restrict__ fix_src)
{
  __shared__ int wsum[kThreads / 64];
  long           dst, src[8];
  int            nsrc = 0, cnt = 0;
  double         ratio = 1.0;
  const long     t = long(blockIdx.x) * kThreads + threadIdx.x;
  const bool     f = slot_needs_fix(m, t, total, dst, src, nsrc, cnt, ratio);
  const unsigned long long mask = __ballot(f);
  const int      lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    wsum[wave] = __popcll(mask);
  __syncthreads();
  if (!f)
    return;
  int pos = offsets[blockIdx.x] + __popcll(mask & ((1ull << lane) - 1ull));
  for (int w = 0; w < wave; ++w)
    pos += wsum[w];
  EmfFix r;
  r.dst = dst;
  r.first = int32_t(pos) * 8;
  r.nsrc = int16_t(nsrc);
  r.cnt = int16_t(cnt);
  r.ratio = ratio;
  fix[pos] = r;
  for (int n = 0; n < nsrc; ++n)
    fix_src[long(pos) * 8 + n] = src[n];
}

} // namespace

int
build_border_list(Context & c)
{
  auto fail = [&](const char * what, hipError_t e) {
    c.err = std::string("border list: ") + what + ": " + hipGetErrorString(e);
    return -1;
  };
  DevMesh &  m = c.dm;
  const long total = long(m.nowned) * border_slots_per_block(m.dim, m.b);
  m.nfix = 0;
  if (total == 0)
    return 0;
  const long nblk = (total + kThreads - 1) / kThreads;
  if (nblk > (1L << 31) - 1)
  {
    c.err = "border list: mesh too large";
    return -1;
  }
  int *      counts = nullptr;
  hipError_t e = hipMalloc(&counts, size_t(nblk) * sizeof(int));
  if (e != hipSuccess)
    return fail("hipMalloc", e);
  hipLaunchKernelGGL(k_border_count, dim3(unsigned(nblk)), kThreads, 0, nullptr, m, total, counts);
  std::vector<int> host(static_cast<size_t>(nblk));
  e = hipMemcpy(host.data(), counts, size_t(nblk) * sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess)
  {
    (void)hipFree(counts);
    return fail("count pass", e);
  }
  long sum = 0;
  for (auto & v : host)
  {
    const long x = v;
    v = int(sum);
    sum += x;
  }
  if (sum * 8 > long(INT32_MAX))
  {
    (void)hipFree(counts);
    c.err = "border list: too many corrected edges";
    return -1;
  }
  e = hipMemcpy(counts, host.data(), size_t(nblk) * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMalloc(&m.fix, size_t(sum > 0 ? sum : 1) * sizeof(EmfFix));
  if (e == hipSuccess)
    e = hipMalloc(&m.fix_src, size_t(sum > 0 ? sum : 1) * 8 * sizeof(int64_t));
  if (e != hipSuccess)
  {
    (void)hipFree(counts);
    return fail("allocation", e);
  }
  c.mesh_bytes += size_t(sum) * (sizeof(EmfFix) + 8 * sizeof(int64_t));
  hipLaunchKernelGGL(k_border_emit, dim3(unsigned(nblk)), kThreads, 0, nullptr, m, total, counts, m.fix, m.fix_src);
  e = hipDeviceSynchronize();
  (void)hipFree(counts);
  if (e != hipSuccess)
    return fail("emit pass", e);
  m.nfix = int(sum);
  return 0;
}

} // namespace mhdct
