// border.hip -- builds the CorrectEmfAtBlockBorder gather list (EmfFix records, 8 source slots
// each) of the mesh in a context on the device, once per mesh: a count pass per workgroup of
// border slots, an exclusive scan of the counts on the device (only the total comes back to the
// host, to size the list), and an emit pass that writes each workgroup's records in slot order.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <string>

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
#pragma unroll
  for (int n = 0; n < 8; ++n)
    if (n < nsrc)
      fix_src[long(pos) * 8 + n] = src[n];
}

constexpr int kScanThreads = 1024;

// exclusive scan of counts[0..n) in place, the sum to counts[n] (one workgroup: each thread
// sums a run of counts, the run sums are scanned in LDS, then each thread writes its run)
__global__ void __launch_bounds__(kScanThreads)
k_border_scan(int * __restrict__ counts, long n)
{
  __shared__ long long part[kScanThreads];
  const long           per = (n + kScanThreads - 1) / kScanThreads;
  const long           lo = min(n, long(threadIdx.x) * per), hi = min(n, lo + per);
  long long            s = 0;
  for (long i = lo; i < hi; ++i)
    s += counts[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1)
  {
    const long long v = threadIdx.x >= unsigned(off) ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  long long run = part[threadIdx.x] - s;
  for (long i = lo; i < hi; ++i)
  {
    const int x = counts[i];
    counts[i] = int(run < INT_MAX ? run : INT_MAX);
    run += x;
  }
  if (threadIdx.x == kScanThreads - 1)
    counts[n] = int(run < INT_MAX ? run : INT_MAX);
}

// p holds at least `bytes`: reallocated (with half again as much room) only when too small
template <class T>
hipError_t
grow(T *& p, size_t & cap, size_t bytes)
{
  if (p && bytes <= cap)
    return hipSuccess;
  if (p)
    (void)hipFree(p);
  p = nullptr;
  cap = 0;
  const size_t want = bytes + bytes / 2;
  const hipError_t e = hipMalloc(&p, want);
  if (e == hipSuccess)
    cap = want;
  return e;
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
  hipError_t e = grow(c.border_scan, c.table_cap[8], size_t(nblk + 1) * sizeof(int));
  if (e != hipSuccess)
    return fail("hipMalloc", e);
  int * counts = c.border_scan;
  hipLaunchKernelGGL(k_border_count, dim3(unsigned(nblk)), kThreads, 0, nullptr, m, total, counts);
  hipLaunchKernelGGL(k_border_scan, dim3(1), kScanThreads, 0, nullptr, counts, nblk);
  int sum = 0;
  e = hipMemcpy(&sum, counts + nblk, sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess)
    return fail("count pass", e);
  // the scan saturates at INT_MAX; the sources are addressed as int32 (8 per record)
  if (long(sum) * 8 > long(INT32_MAX))
  {
    c.err = "border list: too many corrected edges";
    return -1;
  }
  const size_t n = size_t(sum > 0 ? sum : 1);
  e = grow(m.fix, c.table_cap[6], n * sizeof(EmfFix));
  if (e == hipSuccess)
    e = grow(m.fix_src, c.table_cap[7], n * 8 * sizeof(int64_t));
  if (e != hipSuccess)
    return fail("allocation", e);
  c.mesh_bytes += c.table_cap[6] + c.table_cap[7] + c.table_cap[8];
  hipLaunchKernelGGL(k_border_emit, dim3(unsigned(nblk)), kThreads, 0, nullptr, m, total, counts, m.fix, m.fix_src);
  e = hipDeviceSynchronize();
  if (e != hipSuccess)
    return fail("emit pass", e);
  m.nfix = sum;
  return 0;
}

} // namespace mhdct
