// handles.hpp -- the opaque C-ABI handle types of include/mhdct.h.
#pragma once

#include <string>
#include <vector>

#include "amr.hpp"
#include "config.hpp"
#include "ctx.hpp"

struct mhdct_mesh
{
  mhdct::Mesh m;
};

struct mhdct_config
{
  mhdct::Config c;
};

struct mhdct_ctx
{
  mhdct::Context c;
};

struct mhdct_remap
{
  int                             dim = 3;
  int                             b = 16;
  int                             counts[3] = { 0, 0, 0 };
  std::vector<mhdct::RemapEntry> plan;          // per new owned leaf
  mhdct::RemapLists               lists;         // copy pairs, moved, restricted, prolonged leaves
  mhdct::RemapEntry *             dev = nullptr; // the plan in device memory, on first apply
  int *                           dev_lists = nullptr;
  int                             device = -1;
  std::string                     err;
};
