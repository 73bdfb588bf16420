// fused.hpp -- work-list description of the fused form: one slot per mesh, a batch of slots.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ctx.hpp"

namespace mhdct
{

// One mesh ("slot") of a batch: its tables, its state pointers and its scratch.
template <typename T>
struct FusedSlot
{
  DevMesh        m;
  Gas<T>         gt;  // bulk settings
  Gas<double>    gd;  // primitive recovery
  double         dt;
  double         scaling;
  int            abort_eint;
  const double * U;
  double *       U2;
  const double * B;
  double *       B2;
  T *            Qp;  // [block][5][(b+4)^d]       primitive, 2 ghost layers
  T *            Bg;  // [block][ghosted faces]     face field, 2 ghost layers
  T *            Q2;  // [block][8][(b+2)^d]       predictor
  T *            S;   // [block][dim][8][(b+2)^d]  limited slopes
  T *            sF;  // [block][dim][(b+2)^d]     face-B half-step source
  T *            FS;  // [block][6|2][(b+2)^d]     transverse slopes of the face field
  T *            BF;  // [block][dim][2][6][b^(d-1)] border-face fluxes
  T *            emf; // [block][3|1][(b+1)^d]
  T *            Ed;  // [block][3|1][(b+3)^d]     edge electric field v x B at t_n
  const int *    bside; // sides (block*2*dim + 2*dir + side) whose face neighbour is coarser
  int            nbside;
  long           cnt[7]; // work items per stage (the stage offsets are padded to workgroup multiples)
};

// Scratch element counts per block, for a block edge b.
struct FusedSizes
{
  long qp, bg, q2, s, sf, fs, bf, emf, ed;
  FusedSizes(int dim, int b)
  {
    const long g2 = dim == 3 ? long(b + 4) * (b + 4) * (b + 4) : long(b + 4) * (b + 4);
    const long g1 = dim == 3 ? long(b + 2) * (b + 2) * (b + 2) : long(b + 2) * (b + 2);
    const long n = b + 4;
    bg = dim == 3 ? 3 * (n + 1) * n * n : 2 * (n + 1) * n + n * n;
    qp = 5 * g2;
    q2 = 8 * g1;
    s = long(dim) * 8 * g1;
    sf = long(dim) * g1;
    fs = (dim == 3 ? 6 : 2) * g1;
    bf = long(dim) * 2 * 6 * (dim == 3 ? long(b) * b : long(b));
    emf = dim == 3 ? 3L * (b + 1) * (b + 1) * (b + 1) : long(b + 1) * (b + 1);
    ed = dim == 3 ? 3L * (b + 3) * (b + 3) * (b + 3) : long(b + 3) * (b + 3);
  }
  long total() const { return qp + bg + q2 + s + sf + fs + bf + emf + ed; }
};

// The device work list of one size class (dimension, block size): per stage, the prefix
// offsets of every slot's work items.
//   stage 0 k_prim (interior cells, then the ghost shell padded to 64), 1 k_cell (ghost-1 cells),
//   2 k_border_flux (faces of the sides in bside, each side padded to 64), 3 k_flux_tile (8^dim tiles), 4 k_edge (edges, each direction
//   padded to 64), 5 k_bupdate (faces), 6 k_efield (edge points of the E grid, padded to 64)
template <typename T>
struct FusedBatch
{
  int              dim = 3;
  int              bs = 16;
  int              nslots = 0;
  FusedSlot<T> *   d_slots = nullptr;
  static constexpr int kStages = 7;
  long *           d_off[kStages] = {};
  long             tot[kStages] = {};
  int              nfix = 0;
  EmfFix *         fix = nullptr;     // concatenated over slots, dst shifted to the batch EMF array
  int64_t *        fix_src = nullptr;
  T *              fixval = nullptr;
  T *              emf_all = nullptr; // start of the batch EMF array (slot arrays are contiguous in it)
  unsigned long long * flags = nullptr;
  int              grid_cap = 0;      // persistent grid: at most this many workgroups per launch
};

// block sizes the fused kernels are instantiated for
bool fused_supports(int dim, int b);
// work items of one block in stage `stage` (the launch geometry of the kernels above)
long fused_items_per_block(int dim, int b, int stage);
// work items of one coarse-facing side in the border-flux stage
long fused_border_items_per_side(int dim, int b);

template <typename T>
int fused_launch(FusedBatch<T> & bt, hipStream_t st);

} // namespace mhdct
