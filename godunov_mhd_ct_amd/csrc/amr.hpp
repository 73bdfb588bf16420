// amr.hpp -- the device side of the AMR cycle: refinement flags and the remap of U and face B
// between two leaf sets (amr.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mhdct.h"
#include "ctx.hpp"

namespace mhdct
{

// [amr] keys of the refinement criterion (SolverGodunovMHD.cpp:1107-1113)
struct AmrParams
{
  int    level_min;
  int    level_max;
  double epsilon_refine;
  double epsilon_coarsen;
  double epsilon_lohner;
  int    ivar;      // the single refinement variable when nvars == 0
  int    criterion; // 0 LOHNER_UNSPLIT, 1 SIMPLE_GRADIENT
  int    nvars;     // refinement variables in vars (0: ivar alone)
  int    vars[4];   // 0-4 a conserved variable of U, 10-17 a derived quantity (mhdct.h MHDCT_RVAR_*)
};

enum
{
  kRemapCopy = 0,
  kRemapProlong = 1,
  kRemapRestrict = 2
};

// how one new owned leaf gets its data: a copy of an old leaf, a prolongation from its old
// parent (child octant `oct`, x bit 0, y bit 1, z bit 2), or a restriction of its 2^dim old
// children (src in octant order).  nb[2c + side]: the old leaf at the new leaf's level across
// its low (side 0) or high (side 1) face plane along c, or -1; its faces on that plane are
// taken as they are, so that a face plane shared by two leaves of one level keeps one value.
struct RemapEntry
{
  int kind = kRemapCopy;
  int oct = 0;
  int src[8] = { -1, -1, -1, -1, -1, -1, -1, -1 };
  int nb[6] = { -1, -1, -1, -1, -1, -1 };
};

// The checks of the refinement entry points (-1 with c.err set), and the parameters in device form
int check_amr_params(Context & c, const double * U, const double * B, const mhdct_amr_params * p, AmrParams & out);
// Indicator and flags of every owned block into host arrays (flags: +1 refine, -1 coarsen,
// 0 keep; ind optional).  Synchronises `s`.
int launch_refine_flags(Context & c, const double * U, const double * B, const AmrParams & p, int8_t * flags_host,
                        double * ind_host, hipStream_t s);
// The same flags with no device allocation: the buffers are a file-scope device array for at
// most MHDCT_NOALLOC_MAX_BLOCKS owned blocks (-8 beyond it); calls that overlap on one GPU race.
int launch_refine_flags_noalloc(Context & c, const double * U, const double * B, const AmrParams & p,
                                int8_t * flags_host, double * ind_host, hipStream_t s);
// The work lists of a remap, in one int array: copy pairs (old leaf, new leaf), then the new
// leaves that are prolonged or restricted ("moved"), then the restricted ones, then the
// prolonged ones.
struct RemapLists
{
  std::vector<int> ints;
  int              ncopy = 0, nmoved = 0, nrest = 0, nprol = 0;
  const int *      copies(const int * base) const { return base; }
  const int *      moved(const int * base) const { return base + 2 * ncopy; }
  const int *      restricted(const int * base) const { return base + 2 * ncopy + nmoved; }
  const int *      prolonged(const int * base) const { return base + 2 * ncopy + nmoved + nrest; }
};

// Host plan for the owned leaves of new_mesh from those of old_mesh; counts = copied,
// prolonged, restricted.
bool remap_plan(const Mesh & old_mesh, const Mesh & new_mesh, std::vector<RemapEntry> & plan, RemapLists & lists,
                int counts[3], std::string & err);
// lists_dev: the device copy of lists.ints
int launch_remap(int dim, int b, const RemapEntry * plan_dev, const RemapLists & lists, const int * lists_dev,
                 const double * Uo, const double * Bo, double * Un, double * Bn, hipStream_t s, std::string & err);

} // namespace mhdct
