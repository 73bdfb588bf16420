// fused_host.hpp -- the fused form's per-context entry points (fused_host.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "ctx.hpp"
#include "fused.hpp"

namespace mhdct
{

// allocate the context's slot (scratch, work list) for the mesh just uploaded (its host
// neighbour table gives the coarse-facing sides); 0 or <0
int  fused_alloc(Context & c, const Mesh & mesh);
void fused_free(Context & c);
// one step with the caller's dt: U2, B2 <- step(U, B)
int  fused_step(Context & c, const double * U, double * U2, const double * B, double * B2, double dt,
                hipStream_t s);
// the context's slot (host copy), or null if the context is not of that precision
template <typename T>
const FusedSlot<T> * fused_slot(const Context & c);

} // namespace mhdct
