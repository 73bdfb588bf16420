// fused.hpp -- work-list description of the fused form: one slot per mesh, a batch of slots.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "ctx.hpp"

namespace mhdct
{

// Stages of one fused step, in launch order: prim, efield, cell, border, tile, edge, fix
// (k_fix + k_scatter), bupdate; the dt reduction is a separate launch before them.
enum : int
{
  kStagePrim = 0,    // interior cells, then the ghost shell of the 2-ghost tile (padded to 64)
  kStageCell = 1,    // ghost-1 cells (padded to 64 per block)
  kStageBorder = 2,  // faces of the coarse-facing block sides (padded to 64 per side)
  kStageTile = 3,    // 8^dim flux tiles (one workgroup each)
  kStageEdge = 4,    // k_edge wavefronts of a block (fused_edge_waves, x 64)
  kStageBupdate = 5, // faces
  kStageEfield = 6,  // edge points of the E grid (padded to 64 per block)
  kStageFix = 7,     // corrected border edges (the mesh's gather list)
  kStageInvDt = 8,   // owned cells (dt reduction)
  kStageFront = 9,   // sub-tiles of k_front (8 x 8 x fused_front_depth cells)
  kStages = 10
};

// depth along z of a k_front sub-tile (3D: 2 in 8^3 blocks, else 4; 2D: the 8 x 8 tile)
constexpr int
fused_front_depth(int dim, int b)
{
  return dim == 3 ? (b == 8 ? 2 : 4) : 1;
}

// k_cell in the FP64 3D form: the runs along z of each (i, j) column of a block's 1-ghost tile
// (b + 2 cells), each run marched by one lane: runs of 5, 3 and 2 cells for 8^3, 16^3 and 32^3
// blocks, measured on MI355X against the other run lengths (16^3: 2, 6 and 9 cells; 32^3: 17) and
// against one cell per item at 3 wavefronts per SIMD, which was slower at 16^3 and 32^3 (8^3:
// within the noise).  The FP32 form keeps one cell per item (its march at 2 wavefronts per SIMD:
// BL3D 1.06 -> 1.26 ms).  0: one cell per item.
constexpr int
fused_cell_runs(int dim, int b, bool fp64)
{
  return (dim != 3 || !fp64) ? 0 : (b == 8 ? 2 : (b == 16 ? 6 : 17));
}

// Wavefronts of k_edge per block.  In 3D the edges march along z, one (i, j) position per lane:
// b (b + 1) x-edges, (b + 1) b y-edges and (b + 1)^2 z-edges per plane.  In 2D the (b+1)^2
// edges one per lane.
constexpr int
fused_edge_waves(int dim, int b)
{
  const int n1 = b + 1;
  if (dim == 2)
    return (n1 * n1 + 63) / 64;
  const int wp = (n1 * b + 63) / 64;
  const int wx = (b * n1 + 63) / 64;
  const int wz = (n1 * n1 + 63) / 64;
  return wx + wp + wz;
}

// One mesh ("slot") of a batch: its tables, its state pointers and its scratch.
template <typename T>
struct FusedSlot
{
  DevMesh              m;
  Gas<T>               gt; // bulk settings
  Gas<double>          gd; // primitive recovery and dt
  double               dt;
  double               cfl;
  double               scaling;
  int                  abort_eint;
  int                  rk2; // RK2_SSP stage: no half-step predictor, no face-B source at the edges
  double               visc_mu; // >= 0: the HANCOCK predictor's viscous term with this mu; < 0: none
  const double *       U;
  double *             U2;
  const double *       B;
  double *             B2;
  T *                  Qp;     // [block][5][(b+4)^d]       primitive, 2 ghost layers
  T *                  Bg;     // [block][ghosted faces]     face field, 2 ghost layers
  T *                  Q2;     // [block][8][(b+2)^d]       predictor
  T *                  S;      // [block][dim][8][(b+2)^d]  limited slopes
  T *                  sF;     // [block][dim][(b+2)^d]     face-B half-step source
  T *                  FS;     // [block][6|2][(b+2)^d]     transverse slopes of the face field
  T *                  BF;     // [block][dim][2][6][b^(d-1)] border-face fluxes
  T *                  emf;    // [block][3|1][(b+1)^d]
  T *                  Ed;     // [block][3|1][(b+3)^d]     edge electric field v x B at t_n
  T *                  fixval; // [nfix]                    staged corrected border EMFs
  const int *          bside;  // sides (block*2*dim + 2*dir + side) whose face neighbour is coarser
  int                  nbside;
  unsigned long long * red; // [2]: inv_dt bits (max), error flags (bit 0 eint < 0, bit 1 NaN speed)
  int                  dt_red; // 1: k_prim reduces the inverse time step of the slot's owned cells into red[0]
  long                 cnt[kStages]; // work items per stage
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

// Launch tunables of a batch (picked by the autotuner or set by hand).
struct FusedTune
{
  int grid_cap = 1 << 20; // element kernels: at most this many workgroups (multiple of 8)
  int tile_cap = 1 << 20; // flux tiles: at most this many workgroups (multiple of 8)
  int tile_z = 8;         // 3D flux tile depth: 8 (8x8x8 cells, 576 threads) or 4 (8x8x4, 320 threads)
  int flux_march = 1;     // FP64 3D, depth 8: k_flux_march (each record read once per direction), else k_flux_tile
  int front = 0;          // 1: k_front runs the prim, efield and cell stages (kStagePrim; the other
                          // two launch nothing); 0: k_prim, k_efield and k_cell (viscosity needs them)
  int cell_threads = 256; // workgroup size of k_cell: 256 or 512
  int edge_threads = 256; // workgroup size of k_edge: 256 or 512
  int per_class = 1;      // slots of several block sizes: 1 launches every stage once per block-size
                          // class with the kernels specialised to it; 0 once with the kernels that
                          // dispatch on the slot's block size
};

// Block-size classes of a batch of mixed block sizes: the slots are ordered by block size, class c
// holds the slots [s0, s0 + n) of block size bs, with work lists of their own.
struct FusedClass
{
  int    bs = 0, s0 = 0, n = 0;
  long * d_off[kStages] = {};
  long   tot[kStages] = {};
};

// The device work list of a batch: per stage, the prefix offsets of every slot's work items
// (each slot's range padded to 512), and the slot array in two parities: parity 0 advances
// U -> U2, parity 1 advances U2 -> U.  RK2_SSP batches use these two arrays for the first stage
// (into the slots' stage buffers) and arrays 2 and 3 for the second (from them).
template <typename T>
struct FusedBatch
{
  int            dim = 3;
  int            bsel = 0; // the block size of every slot, or 0 when they differ
  int            rk2 = 0;  // every slot is an RK2_SSP stage (a batch holds one integrator)
  int            grav = 0; // some slot has gravity
  int            dev_dt = 0; // dt on the device (a batch): reduced in k_prim and set before k_efield
  int            visc = 0;   // the slot has the viscous predictor term (per-mesh FP64 form only)
  int            nslots = 0;
  FusedSlot<T> * d_slots[4] = { nullptr, nullptr, nullptr, nullptr };
  long *         d_off[kStages] = {};
  long           tot[kStages] = {};
  double *       clock = nullptr; // [nslots][2]: simulated time, steps (device-side dt only)
  FusedTune      tune;
  int            ncls = 0; // block-size classes (bsel 0 only)
  FusedClass     cls[3];
};

// block sizes the fused kernels are instantiated for
bool fused_supports(int dim, int b);
// work items of one block in stage `stage` (the launch geometry of the kernels above)
long fused_items_per_block(int dim, int b, int stage, bool fp64);
// work items of one coarse-facing side in the border-flux stage
long fused_border_items_per_side(int dim, int b);
// a slot's work range, padded so that every workgroup and wavefront lies in one slot
inline long
fused_pad_slot(long n)
{
  return (n + 511) / 512 * 512;
}

// the fused stages of one step for the slot array of `parity` (dt already in the slots)
template <typename T>
int fused_launch(const FusedBatch<T> & bt, int parity, hipStream_t st);
// only stage `which` (kStagePrim ... kStageBupdate; kStageFix launches k_fix and k_scatter)
template <typename T>
int fused_launch_stage(const FusedBatch<T> & bt, int parity, int which, hipStream_t st);
// the launch order of the stages of one step
constexpr int fused_stage_order[8] = { kStagePrim, kStageEfield, kStageCell, kStageBorder,
                                       kStageTile, kStageEdge, kStageFix,
                                       kStageBupdate };
// dt on the device: clear the slots' reduction words, reduce over the owned cells of the state
// of `parity`, then dt = cfl / inv_dt into every slot array and the clock advanced
template <typename T>
int fused_launch_dt(const FusedBatch<T> & bt, int parity, hipStream_t st);

} // namespace mhdct
