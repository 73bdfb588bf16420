// fused.hip -- fused per-block form (placeholder until the fused kernels land).
#include <hip/hip_runtime.h>

#include "ctx.hpp"

namespace mhdct
{

int
fused_alloc(Context & c)
{
  c.err = "fused form not built yet";
  return -4;
}

void
fused_free(Context & c)
{
  (void)c;
}

int
fused_step(Context & c, const double *, double *, const double *, double *, double, hipStream_t)
{
  c.err = "fused form not built yet";
  return -4;
}

} // namespace mhdct
