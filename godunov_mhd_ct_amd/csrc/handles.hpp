// handles.hpp -- the opaque C-ABI handle types of include/mhdct.h.
#pragma once

#include "ctx.hpp"

struct mhdct_mesh
{
  mhdct::Mesh m;
};

struct mhdct_ctx
{
  mhdct::Context c;
};
