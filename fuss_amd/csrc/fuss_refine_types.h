// fuss_refine_types.h -- per-block descriptor of the correction multigrid
// levels (fuss_refine), shared by the host context and the kernels.
#pragma once

namespace fuss {

struct CBlk {
  int n[3];     // cells of the block at this level
  int r[3];     // aggregation ratio from the finer level per direction (1 or 2)
  int nf[3];    // cells of the block at the finer level
  int off;      // first cell of the block in this level's arrays
  int foff;     // first cell of the block at the finer level
  int pad_;
};

}  // namespace fuss
