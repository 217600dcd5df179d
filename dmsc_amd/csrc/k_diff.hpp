// Persistence-loss kernels (k_diff.hip): descriptors of a chunk of fields, passed to the kernels by
// value, and their launchers.
#pragma once

#include "internal.hpp"

namespace dmsc {

constexpr int kBwdChunk = 32;  // fields per launch (the descriptors travel as kernel arguments)

struct BwdField {
  const float* f;      // H x W field (backward only)
  float* grad;         // H x W gradient, fully overwritten (backward only)
  const float* gscale; // device scalar the field's gradient is scaled by (backward only)
  const int32_t* sad_pts;
  const float* p_min;
  const int32_t* pp_min;
  const int32_t* min_pts;
  const float* p_max;
  const int32_t* pp_max;
  const int32_t* max_pts;
  int32_t H, W, n_sad, n_min, n_max;
  int32_t blk_sad, blk_pix;  // first block of this field in the scatter / gather launches
  int64_t scratch_base;      // first scratch float of this field (bwd_scratch_floats each)
};

struct BwdChunk {
  BwdField fld[kBwdChunk];
  int32_t n;  // fields in this chunk
  int32_t blocks_sad, blocks_pix;
  int64_t scratch;    // scratch floats the chunk uses
  int64_t loss_base;  // index of the chunk's first field in the loss output
};

// scratch of one H x W field: nine planes of H W floats (the cells around each value pixel)
inline int64_t bwd_scratch_floats(int H, int W) { return 9 * (int64_t)H * W; }
// block bases and scratch offsets of the chunk's fields (host side)
void bwd_assign_blocks(BwdChunk& c);
// most blocks one field may need in a launch (the chunking keeps a launch's grid within int32)
int64_t bwd_blocks_for(int H, int W, int64_t n_sad);
// scatter and gather of one chunk; `scratch` holds at least c.scratch floats, all zero before and after
void launch_persistence_backward_chunk(bool dual, const BwdChunk& c, float max_pers, float* scratch, hipStream_t st);
// partial sums the loss of a chunk of n fields needs (doubles)
int64_t loss_partials(int n_fields);
// loss[c.loss_base + k] for each field k of the chunk; `part` holds loss_partials(c.n) doubles
void launch_persistence_loss_chunk(const BwdChunk& c, float max_pers, double* part, float* loss, hipStream_t st);

}  // namespace dmsc
