// wf_octree.h -- host launchers of the level-synchronous octree build.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wf {

int64_t scan_tmp_ints(int64_t n);
int64_t sort_count_ints(int64_t n);

// Exclusive scan of int32 (in may equal out).  tmp: scan_tmp_ints(n) ints.
// total_dev (nullable) receives the sum.
hipError_t exclusive_scan(const int32_t* in, int32_t* out, int64_t n, int32_t* tmp,
                          int32_t* total_dev, hipStream_t st);

// Keys from the per-axis float split-plane tables (planes_dev: 3 x (2^depth - 1)
// floats, level l cell i at (2^l - 1) + i) and a stable LSD radix sort, one
// octant digit per pass.  Results in keys/idx (idx = original point index).
hipError_t octree_sort(const float* x, const float* y, const float* z, int64_t n,
                       const float* planes_dev, int depth, uint32_t* keys, int32_t* idx,
                       uint32_t* keys_tmp, int32_t* idx_tmp, int32_t* counts, int32_t* scan_tmp,
                       hipStream_t st);

// Active (non-empty) nodes of one level: node ids act_key (= key >> shift) and
// first point act_begin, in level order; count_dev receives the count.
hipError_t octree_level(const uint32_t* keys, int64_t n, int shift, int32_t* flags, int32_t* pos,
                        int32_t* scan_tmp, int32_t* count_dev, uint32_t* act_key,
                        int32_t* act_begin, hipStream_t st);
hipError_t octree_level_finish(const int32_t* act_begin, int count, int32_t n, int32_t* act_end,
                               hipStream_t st);
// Child-level -> parent-level links: parent_idx per child, [cb, ce) per parent.
hipError_t octree_links(const int32_t* child_begin, int nchild, const int32_t* parent_begin,
                        int nparent, int32_t* parent_idx, int32_t* cb, int32_t* ce,
                        hipStream_t st);
// Ranges of all (8^(depth+1)-1)/7 nodes in level order.
hipError_t octree_node_ranges(const uint32_t* keys, int64_t n, int depth, int32_t* nb,
                              int32_t* ne, hipStream_t st);
hipError_t gather_points(const int32_t* perm, int64_t n, const float* const src[7], float* out7,
                         hipStream_t st);

}  // namespace wf
