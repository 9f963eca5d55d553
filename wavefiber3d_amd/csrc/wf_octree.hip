// wf_octree.hip -- level-synchronous octree build.
//
// Replaces the recursive CUDA-dynamic-parallelism build_octree_kernel
// (util/tree.h:289-715, (8^d-1)/7 device launches of one 256-thread block per
// node) and the host passes of treePO.cu:256-337.  The reference partitions
// each node stably into 8 octants (tree.h:492-625) with the float centre
// 0.5f*(min+max) of the node box (tree.h:116-120) and octant code
// 4*[z<cz] + 2*[y<cy] + [x>=cx] (predicates tree.h:379-414, child boxes
// tree.h:672-687).  A recursive stable partition by digits equals a stable sort
// by the digit string with the level-1 digit most significant, so:
//   1. each point gets its d-digit key by descending the per-axis tables of
//      float split planes (the reference's exact float centres),
//   2. a stable LSD radix sort with one octant digit (3 bits) per pass,
//   3. per level, run starts of key >> 3(d-l) give the active (non-empty)
//      nodes in level order -- the reference's nonemptynodes lists
//      (treePO.cu:262-307); in a full tree every non-empty node has a
//      non-empty child, so "active" == "non-empty".
// Every pass is a plain streaming kernel over N keys (HBM-bound).
#include "wf_common.h"
#include "wf_octree.h"

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;  // consecutive items per thread (keeps order for stability)
constexpr int kSortTile = kSortThreads * kSortItems;

__global__ __launch_bounds__(256) void keys_kernel(const float* __restrict__ x,
                                                   const float* __restrict__ y,
                                                   const float* __restrict__ z, int64_t n,
                                                   const float* __restrict__ planes, int depth,
                                                   uint32_t* __restrict__ keys,
                                                   int32_t* __restrict__ idx) {
  const int per_axis = (1 << depth) - 1;
  const float* px = planes;
  const float* py = planes + per_axis;
  const float* pz = planes + 2 * per_axis;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float a = x[i], b = y[i], c = z[i];
    uint32_t key = 0;
    int ix = 0, iy = 0, iz = 0;
    for (int l = 0; l < depth; ++l) {
      const int off = (1 << l) - 1;
      const float cx = px[off + ix], cy = py[off + iy], cz = pz[off + iz];
      const int hx = a >= cx, hy = b >= cy, hz = c >= cz;
      const uint32_t o = 4u * (uint32_t)(!hz) + 2u * (uint32_t)(!hy) + (uint32_t)hx;
      key = key * 8u + o;
      ix = 2 * ix + hx;
      iy = 2 * iy + hy;
      iz = 2 * iz + hz;
    }
    keys[i] = key;
    idx[i] = (int32_t)i;
  }
}

// Per-block digit histogram: counts[digit * nblocks + block].
__global__ __launch_bounds__(kSortThreads) void radix_count_kernel(const uint32_t* __restrict__ keys,
                                                                   int64_t n, int shift,
                                                                   int32_t* __restrict__ counts,
                                                                   int nblocks) {
  __shared__ int32_t h[8];
  if (threadIdx.x < 8) h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  int local[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int it = 0; it < kSortItems; ++it) {
    const int64_t i = base + (int64_t)threadIdx.x * kSortItems + it;
    if (i < n) local[(keys[i] >> shift) & 7u]++;
  }
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (local[q]) atomicAdd(&h[q], local[q]);
  __syncthreads();
  if (threadIdx.x < 8) counts[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: thread t owns items [t*I, (t+1)*I) of the tile in order;
// local offsets come from an exclusive scan over (digit, thread) in LDS.
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const uint32_t* __restrict__ kin, const int32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    int32_t* __restrict__ vout, int64_t n, int shift, const int32_t* __restrict__ offs,
    int nblocks) {
  __shared__ int32_t cnt[8][kSortThreads];
  __shared__ int32_t tot[8];
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
  uint32_t kk[kSortItems];
  int32_t vv[kSortItems];
  int local[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int64_t i = base + (int64_t)threadIdx.x * kSortItems + it;
    if (i < n) {
      kk[it] = kin[i];
      vv[it] = vin[i];
      local[(kk[it] >> shift) & 7u]++;
    } else {
      kk[it] = 0xffffffffu;
      vv[it] = -1;
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) cnt[q][threadIdx.x] = local[q];
  __syncthreads();
  // exclusive scan along threads for each digit (8 warps... 4 wave64s: each
  // wave scans two digit rows serially in chunks of 64)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int q = wave; q < 8; q += kSortThreads / 64) {
    int carry = 0;
    for (int c0 = 0; c0 < kSortThreads; c0 += 64) {
      int v = cnt[q][c0 + lane];
      int incl = v;
      for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(incl, off);
        if (lane >= off) incl += t;
      }
      cnt[q][c0 + lane] = carry + incl - v;
      carry += __shfl(incl, 63);
    }
    if (lane == 0) tot[q] = carry;
  }
  __syncthreads();
  int run[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) run[q] = offs[q * nblocks + blockIdx.x] + cnt[q][threadIdx.x];
#pragma unroll
  for (int it = 0; it < kSortItems; ++it) {
    const int64_t i = base + (int64_t)threadIdx.x * kSortItems + it;
    if (i < n) {
      const int q = (kk[it] >> shift) & 7u;
      const int dst = run[q]++;
      kout[dst] = kk[it];
      vout[dst] = vv[it];
    }
  }
}

// ---------------------------------------------------------------- scans --
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline int block_exclusive_scan(int v, int* sh /*kScanThreads/64 + 1*/, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(incl, off);
    if (lane >= off) incl += t;
  }
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) {
      int t = sh[w];
      sh[w] = acc;
      acc += t;
    }
    sh[kScanThreads / 64] = acc;
  }
  __syncthreads();
  int r = sh[wave] + incl - v;
  *total = sh[kScanThreads / 64];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const int32_t* __restrict__ in,
                                                                   int64_t n,
                                                                   int32_t* __restrict__ bsum) {
  __shared__ int sh[kScanThreads / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int s = 0;
  for (int it = 0; it < kScanItems; ++it)
    if (base + it < n) s += in[base + it];
  int total;
  block_exclusive_scan(s, sh, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void scan_sums_kernel(int32_t* __restrict__ bsum,
                                                                 int nb, int32_t* __restrict__ total_out) {
  __shared__ int sh[kScanThreads / 64 + 1];
  int carry = 0;
  for (int c0 = 0; c0 < nb; c0 += kScanThreads) {
    const int i = c0 + threadIdx.x;
    const int v = i < nb ? bsum[i] : 0;
    int total;
    const int e = block_exclusive_scan(v, sh, &total);
    if (i < nb) bsum[i] = carry + e;
    carry += total;
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const int32_t* __restrict__ in,
                                                                  int64_t n,
                                                                  const int32_t* __restrict__ bsum,
                                                                  int32_t* __restrict__ out) {
  __shared__ int sh[kScanThreads / 64 + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
  for (int it = 0; it < kScanItems; ++it) {
    v[it] = (base + it < n) ? in[base + it] : 0;
    s += v[it];
  }
  int total;
  int e = block_exclusive_scan(s, sh, &total) + bsum[blockIdx.x];
  for (int it = 0; it < kScanItems; ++it) {
    if (base + it < n) out[base + it] = e;
    e += v[it];
  }
}

// ----------------------------------------------------------- level lists --
__global__ __launch_bounds__(256) void level_flags_kernel(const uint32_t* __restrict__ keys,
                                                          int64_t n, int shift,
                                                          int32_t* __restrict__ flags) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    flags[i] = (i == 0) || ((keys[i] >> shift) != (keys[i - 1] >> shift));
}

__global__ __launch_bounds__(256) void level_compact_kernel(const uint32_t* __restrict__ keys,
                                                            const int32_t* __restrict__ flags,
                                                            const int32_t* __restrict__ pos,
                                                            int64_t n, int shift,
                                                            uint32_t* __restrict__ act_key,
                                                            int32_t* __restrict__ act_begin) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (flags[i]) {
      act_key[pos[i]] = keys[i] >> shift;
      act_begin[pos[i]] = (int32_t)i;
    }
  }
}

__global__ __launch_bounds__(256) void level_end_kernel(const int32_t* __restrict__ act_begin,
                                                        int count, int32_t n,
                                                        int32_t* __restrict__ act_end) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) act_end[i] = (i + 1 < count) ? act_begin[i + 1] : n;
}

// parent active index of each child-level node: the parent whose range holds
// the child's first point (upper_bound over the parent begins, minus one).
__global__ __launch_bounds__(256) void parent_index_kernel(const int32_t* __restrict__ child_begin,
                                                           int nchild,
                                                           const int32_t* __restrict__ parent_begin,
                                                           int nparent,
                                                           int32_t* __restrict__ parent_idx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nchild) return;
  const int32_t b = child_begin[j];
  int lo = 0, hi = nparent;  // first index with parent_begin > b
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (parent_begin[mid] <= b) lo = mid + 1;
    else hi = mid;
  }
  parent_idx[j] = lo - 1;
}

__global__ __launch_bounds__(256) void child_ranges_kernel(const int32_t* __restrict__ parent_idx,
                                                           int nchild, int nparent,
                                                           int32_t* __restrict__ cb,
                                                           int32_t* __restrict__ ce) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nparent) return;
  int lo = 0, hi = nchild;  // lower_bound(p)
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (parent_idx[mid] < p) lo = mid + 1;
    else hi = mid;
  }
  cb[p] = lo;
  int lo2 = lo, hi2 = nchild;  // upper_bound(p)
  while (lo2 < hi2) {
    const int mid = (lo2 + hi2) >> 1;
    if (parent_idx[mid] <= p) lo2 = mid + 1;
    else hi2 = mid;
  }
  ce[p] = lo2;
}

// Full reference node array (tree.h Octree_node ranges) by binary search of
// every node's key prefix range in the sorted keys.
__global__ __launch_bounds__(256) void node_ranges_kernel(const uint32_t* __restrict__ keys,
                                                          int64_t n, int depth,
                                                          int32_t* __restrict__ nb,
                                                          int32_t* __restrict__ ne,
                                                          int64_t nnodes) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnodes) return;
  int level = 0;
  int64_t off = 0, p8 = 1;
  while (off + p8 <= j) {
    off += p8;
    p8 *= 8;
    ++level;
  }
  const uint64_t m = (uint64_t)(j - off);
  const int sh = 3 * (depth - level);
  const uint64_t lo_key = m << sh, hi_key = (m + 1) << sh;
  auto lower = [&](uint64_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if ((uint64_t)keys[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    return (int32_t)lo;
  };
  nb[j] = lower(lo_key);
  ne[j] = lower(hi_key);
}

__global__ __launch_bounds__(256) void gather_kernel(const int32_t* __restrict__ perm, int64_t n,
                                                     const float* __restrict__ a0,
                                                     const float* __restrict__ a1,
                                                     const float* __restrict__ a2,
                                                     const float* __restrict__ a3,
                                                     const float* __restrict__ a4,
                                                     const float* __restrict__ a5,
                                                     const float* __restrict__ a6,
                                                     float* __restrict__ out /*7 x n*/) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t p = perm[i];
    out[i] = a0[p];
    out[n + i] = a1[p];
    out[2 * n + i] = a2[p];
    out[3 * n + i] = a3[p];
    out[4 * n + i] = a4[p];
    out[5 * n + i] = a5[p];
    out[6 * n + i] = a6[p];
  }
}

inline unsigned grid_for(int64_t n, int threads = 256, int64_t cap = 8192) {
  int64_t b = (n + threads - 1) / threads;
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace

namespace wf {

int64_t scan_tmp_ints(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

hipError_t exclusive_scan(const int32_t* in, int32_t* out, int64_t n, int32_t* tmp,
                          int32_t* total_dev, hipStream_t st) {
  const int nb = (int)((n + kScanTile - 1) / kScanTile);
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(kScanThreads), 0, st, in, n, tmp);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanThreads), 0, st, tmp, nb, total_dev);
  hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(kScanThreads), 0, st, in, n, tmp, out);
  return hipGetLastError();
}

int64_t sort_count_ints(int64_t n) { return 8 * ((n + kSortTile - 1) / kSortTile); }

hipError_t octree_sort(const float* x, const float* y, const float* z, int64_t n,
                       const float* planes_dev, int depth, uint32_t* keys, int32_t* idx,
                       uint32_t* keys_tmp, int32_t* idx_tmp, int32_t* counts, int32_t* scan_tmp,
                       hipStream_t st) {
  hipLaunchKernelGGL(keys_kernel, dim3(grid_for(n)), dim3(256), 0, st, x, y, z, n, planes_dev,
                     depth, keys, idx);
  const int nblocks = (int)((n + kSortTile - 1) / kSortTile);
  uint32_t* ka = keys;
  int32_t* va = idx;
  uint32_t* kb = keys_tmp;
  int32_t* vb = idx_tmp;
  for (int pass = 0; pass < depth; ++pass) {  // least significant digit = deepest level first
    const int shift = 3 * pass;
    hipLaunchKernelGGL(radix_count_kernel, dim3(nblocks), dim3(kSortThreads), 0, st, ka, n, shift,
                       counts, nblocks);
    hipError_t e = exclusive_scan(counts, counts, 8 * (int64_t)nblocks, scan_tmp, nullptr, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(radix_scatter_kernel, dim3(nblocks), dim3(kSortThreads), 0, st, ka, va, kb,
                       vb, n, shift, counts, nblocks);
    uint32_t* tk = ka; ka = kb; kb = tk;
    int32_t* tv = va; va = vb; vb = tv;
  }
  if (ka != keys) {
    hipMemcpyAsync(keys, ka, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st);
    hipMemcpyAsync(idx, va, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st);
  }
  return hipGetLastError();
}

hipError_t octree_level(const uint32_t* keys, int64_t n, int shift, int32_t* flags, int32_t* pos,
                        int32_t* scan_tmp, int32_t* count_dev, uint32_t* act_key,
                        int32_t* act_begin, hipStream_t st) {
  hipLaunchKernelGGL(level_flags_kernel, dim3(grid_for(n)), dim3(256), 0, st, keys, n, shift,
                     flags);
  hipError_t e = exclusive_scan(flags, pos, n, scan_tmp, count_dev, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(level_compact_kernel, dim3(grid_for(n)), dim3(256), 0, st, keys, flags, pos,
                     n, shift, act_key, act_begin);
  return hipGetLastError();
}

hipError_t octree_level_finish(const int32_t* act_begin, int count, int32_t n, int32_t* act_end,
                               hipStream_t st) {
  hipLaunchKernelGGL(level_end_kernel, dim3((count + 255) / 256), dim3(256), 0, st, act_begin,
                     count, n, act_end);
  return hipGetLastError();
}

hipError_t octree_links(const int32_t* child_begin, int nchild, const int32_t* parent_begin,
                        int nparent, int32_t* parent_idx, int32_t* cb, int32_t* ce,
                        hipStream_t st) {
  hipLaunchKernelGGL(parent_index_kernel, dim3((nchild + 255) / 256), dim3(256), 0, st,
                     child_begin, nchild, parent_begin, nparent, parent_idx);
  hipLaunchKernelGGL(child_ranges_kernel, dim3((nparent + 255) / 256), dim3(256), 0, st,
                     parent_idx, nchild, nparent, cb, ce);
  return hipGetLastError();
}

hipError_t octree_node_ranges(const uint32_t* keys, int64_t n, int depth, int32_t* nb,
                              int32_t* ne, hipStream_t st) {
  int64_t nn = 0, p8 = 1;
  for (int l = 0; l <= depth; ++l) {
    nn += p8;
    p8 *= 8;
  }
  hipLaunchKernelGGL(node_ranges_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st,
                     keys, n, depth, nb, ne, nn);
  return hipGetLastError();
}

hipError_t gather_points(const int32_t* perm, int64_t n, const float* const src[7], float* out7,
                         hipStream_t st) {
  hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n)), dim3(256), 0, st, perm, n, src[0], src[1],
                     src[2], src[3], src[4], src[5], src[6], out7);
  return hipGetLastError();
}

}  // namespace wf
