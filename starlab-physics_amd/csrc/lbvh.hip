// Broad phase on gfx950: face / mesh AABBs, Morton codes, Karras radix tree,
// bottom-up refit with atomic arrival flags, stack traversal with the stack
// in LDS.
//
// Two trees share the kernels:
//  * the reference's cell LBVH (framework/collision/lbvh_cell.py): a static
//    Karras tree over the Morton-sorted 32^3 cell grid, refilled every build
//    with the faces bucketed by cell (leaf slot i holds row-major cell i,
//    lbvh_cell.py:601-634);
//  * a per-triangle LBVH: one leaf per face in Morton order of the face-box
//    centres, index tie-break for equal codes (64-bit keys code<<32 | index).
//
// Node storage (HBM): nlo[k] / nhi[k] are float4 = (box min, child_a) and
// (box max, child_b) for internal nodes, (box, first face slot) / (box, last
// face slot) for cell leaves and (box, face id) for triangle leaves, so one
// 32-byte fetch gives the traversal everything it needs at a node.
#include "slx_common.h"

namespace slx {

size_t radix_sort_ws_bytes(int n_max);
int radix_sort_pairs(int n_max, const int* n_dev, uint32_t* keys, int32_t* vals, int key_bits,
                     void* ws, size_t ws_bytes, hipStream_t st);
size_t scan_ws_bytes(int n_max);
int exclusive_scan(int n_max, const int* n_dev, const int* in, int* out, int* total_dev, void* ws,
                   size_t ws_bytes, hipStream_t st);

#ifndef SLX_TRAV_WAVE
#define SLX_TRAV_WAVE 1   // wavefront-coherent triangle traversal (0: one walk per lane)
#endif
#ifndef SLX_TRAV_QUAD
#define SLX_TRAV_QUAD 1   // 4-wide node walk when the tree has its quad array (0: binary walk)
#endif
#ifndef SLX_TRAV_SCALAR
#define SLX_TRAV_SCALAR 1 // a line every lane pops is read with scalar loads (0: vector loads only)
#endif
#ifndef SLX_TRAV_STACK
#define SLX_TRAV_STACK 64
#endif
constexpr int kStack = SLX_TRAV_STACK;   // traversal stack depth (entries), twice the reference's 32
// Shallow stacks of the first-tier triangle walk (LDS per block = depth x 128 x 4
// bytes, which bounds the waves in flight); queries that overflow one are walked
// again with kStack entries (depth: knobs().walk_stack).
constexpr int kTravBS = 128;     // traversal block size

// ----------------------------------------------------------------- AABBs

// (corner_box / face_box, the padded box of a face: slx_common.h)

__global__ void k_face_aabbs(int nf, const int32_t* __restrict__ faces, const float4* __restrict__ x4,
                             float pad, float4* __restrict__ fmin, float4* __restrict__ fmax) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  float3 lo, hi;
  face_box(faces, x4, f, pad, lo, hi);
  fmin[f] = make_float4(lo.x, lo.y, lo.z, __int_as_float(f));
  fmax[f] = make_float4(hi.x, hi.y, hi.z, __int_as_float(f));
}

// computeAABB, meshTaichiWrapper.py:334-347: min/max reduction (order-independent,
// hence exact), stage 1 per block, stage 2 one block adds the padding.
constexpr int kBoxBS = 256;
__global__ void __launch_bounds__(kBoxBS)
k_box_partial(int nv, const float4* __restrict__ x4, float* __restrict__ part) {
  __shared__ float s[6][kBoxBS];
  float lo[3] = {1e5f, 1e5f, 1e5f}, hi[3] = {-1e5f, -1e5f, -1e5f};
  for (int i = blockIdx.x * kBoxBS + threadIdx.x; i < nv; i += gridDim.x * kBoxBS) {
    float4 p = x4[i];
    lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
    hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
  }
  for (int c = 0; c < 3; ++c) { s[c][threadIdx.x] = lo[c]; s[3 + c][threadIdx.x] = hi[c]; }
  __syncthreads();
  for (int o = kBoxBS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = fminf(s[c][threadIdx.x], s[c][threadIdx.x + o]);
        s[3 + c][threadIdx.x] = fmaxf(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + o]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void k_box_final(int nparts, const float* __restrict__ part, float pad, float* __restrict__ box) {
  if (threadIdx.x != 0) return;
  float lo[3] = {1e5f, 1e5f, 1e5f}, hi[3] = {-1e5f, -1e5f, -1e5f};
  for (int b = 0; b < nparts; ++b)
    for (int c = 0; c < 3; ++c) {
      lo[c] = fminf(lo[c], part[6 * b + c]);
      hi[c] = fmaxf(hi[c], part[6 * b + 3 + c]);
    }
  for (int c = 0; c < 3; ++c) { box[c] = lo[c] - pad * 1.0f; box[3 + c] = hi[c] + pad * 1.0f; }
}

constexpr int kBoxParts = 1024;

// ----------------------------------------------------------------- Morton

// expand_bits / morton_3d, lbvh_cell.py:85-102 (x in the lowest bit)
__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ float clampf(float a, float lo, float hi) { return fminf(fmaxf(a, lo), hi); }
__device__ __forceinline__ uint32_t morton3d(float x, float y, float z) {
  x = clampf(x * 1024.0f, 0.0f, 1023.0f);
  y = clampf(y * 1024.0f, 0.0f, 1023.0f);
  z = clampf(z * 1024.0f, 0.0f, 1023.0f);
  return expand_bits((uint32_t)x) | (expand_bits((uint32_t)y) << 1) | (expand_bits((uint32_t)z) << 2);
}
// normalised code of a point in [lo, hi] (lbvh_cell.py:139-148)
__device__ __forceinline__ uint32_t morton_in_box(float3 p, float3 lo, float3 hi) {
  return morton3d(clampf((p.x - lo.x) / (hi.x - lo.x), 0.f, 1.f), clampf((p.y - lo.y) / (hi.y - lo.y), 0.f, 1.f),
                  clampf((p.z - lo.z) / (hi.z - lo.z), 0.f, 1.f));
}

// The triangle trees' code of a face centre: normalised by the cube of the box's
// largest extent (the same scale on every axis) instead of each axis's own extent.
// A draped sheet 1 m wide is ~0.1 m tall; per-axis codes would spend as many bits
// on its thickness as on its width and cut its tree along the height, so that
// node boxes of nearby faces come out larger.  On the 1M-vertex drape a cloth-tree
// walk visits 16 lines per query instead of 20 (cache build 297 -> 250 us).  (The
// cell tree keeps its fixed per-axis grid, lbvh_cell.py:122-152; the reference's
// per-triangle Morton pass, lbvh_cell.py:153-194, is never called.)
__device__ __forceinline__ uint32_t morton_in_cube(float3 p, float3 lo, float3 hi) {
  const float e = fmaxf(fmaxf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
  return morton_in_box(p, lo, make_float3(lo.x + e, lo.y + e, lo.z + e));
}

__device__ __forceinline__ float3 cell_size_of(const float* box) {
  return make_float3((box[3] - box[0]) / (float)kGrid, (box[4] - box[1]) / (float)kGrid,
                     (box[5] - box[2]) / (float)kGrid);
}

// lbvh_cell.py:564-574 (truncation, int32 wrap-around flattening)
__device__ __forceinline__ int32_t flat_cell_id(float3 p, float3 cs, float3 o) {
  uint32_t i = (uint32_t)trunc_i32((p.x - o.x) / cs.x), j = (uint32_t)trunc_i32((p.y - o.y) / cs.y),
           k = (uint32_t)trunc_i32((p.z - o.z) / cs.z);
  return (int32_t)(i * (uint32_t)(kGrid * kGrid) + j * (uint32_t)kGrid + k);
}

// assign_cell_morton over [lo, hi], lbvh_cell.py:122-152
__global__ void k_cell_morton(float3 lo, float3 hi, uint32_t* __restrict__ codes, int32_t* __restrict__ ids) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kCells) return;
  float3 cs = make_float3((hi.x - lo.x) / (float)kGrid, (hi.y - lo.y) / (float)kGrid, (hi.z - lo.z) / (float)kGrid);
  int x0 = i / (kGrid * kGrid), y0 = (i % (kGrid * kGrid)) / kGrid, z0 = i % kGrid;
  float3 c = make_float3(((cs.x * (float)x0) + (0.5f * cs.x)) + lo.x, ((cs.y * (float)y0) + (0.5f * cs.y)) + lo.y,
                         ((cs.z * (float)z0) + (0.5f * cs.z)) + lo.z);
  codes[i] = morton_in_box(c, lo, hi);
  ids[i] = i;
}

__global__ void k_face_morton(int nf, const float4* __restrict__ fmin, const float4* __restrict__ fmax,
                              const float* __restrict__ box, uint32_t* __restrict__ codes,
                              int32_t* __restrict__ ids) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  float4 a = fmin[f], b = fmax[f];
  float3 c = make_float3(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z));
  codes[f] = morton_in_cube(c, make_float3(box[0], box[1], box[2]), make_float3(box[3], box[4], box[5]));
  ids[f] = f;
}

// ----------------------------------------------------------------- Karras

// common prefix length of sorted keys i and j; -1 outside [0, n) (lbvh_cell.py:255-264)
__device__ __forceinline__ int prefix_len(const uint32_t* codes, int n, bool wide, int i, int j) {
  if (j < 0 || j >= n) return -1;
  if (wide) {
    uint64_t a = ((uint64_t)codes[i] << 32) | (uint32_t)i, b = ((uint64_t)codes[j] << 32) | (uint32_t)j;
    return __clzll(a ^ b);
  }
  uint32_t x = codes[i] ^ codes[j];
  return x ? __clz(x) : 32;
}

// Karras 2012 internal node i (lbvh_cell.py:301-336, 272-298, 407-430)
__global__ void k_karras(int n, const uint32_t* __restrict__ codes, int wide, int32_t* __restrict__ ca,
                         int32_t* __restrict__ cb, int32_t* __restrict__ parent, int32_t* __restrict__ rl,
                         int32_t* __restrict__ rr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) parent[0] = -1;
  if (i >= n - 1) return;
  const bool W = wide != 0;
  int dl = prefix_len(codes, n, W, i, i - 1), dr = prefix_len(codes, n, W, i, i + 1);
  int d = (dr < dl) ? -1 : 1;
  int dmin = (dr < dl) ? dr : dl;
  int hi = 2;
  while (prefix_len(codes, n, W, i, i + hi * d) > dmin) hi *= 2;
  int len = 0;
  for (int t = hi / 2; t > 0; t /= 2)
    if (prefix_len(codes, n, W, i, i + (len + t) * d) > dmin) len += t;
  int j = i + len * d;
  int first = d > 0 ? i : j, last = d > 0 ? j : i;
  int node_prefix = prefix_len(codes, n, W, first, last);
  int split;
  if (!W && node_prefix == 32) {
    split = (first + last) / 2;
  } else {
    split = first;
    int step = last - first;
    do {
      step = (step + 1) / 2;
      int c = split + step;
      if (c < last && prefix_len(codes, n, W, first, c) > node_prefix) split = c;
    } while (step > 1);
  }
  int a = (split == first) ? split + (n - 1) : split;
  int b = (split + 1 == last) ? split + 1 + (n - 1) : split + 1;
  ca[i] = a;
  cb[i] = b;
  parent[a] = i;
  parent[b] = i;
  if (rl) rl[i] = first;
  if (rr) rr[i] = last;
}

// ----------------------------------------------------------------- refit

__global__ void k_fill_i32(int n, int32_t* __restrict__ p, int32_t v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Bottom-up refit without device-wide synchronisation.
//
// The reference refits with one thread per leaf walking up and atomic arrival
// counters in global memory (lbvh_cell.py:490-522).  On a GPU whose L2 is split
// over several dies every step of such a walk needs a device-scope release and
// acquire, i.e. an L2 write-back and invalidate.  Here a schedule, made once per
// topology, cuts the tree into maximal subtrees of at most T leaves (a leaf range
// of an LBVH node is contiguous, and every internal node's index lies inside its
// own range), each refit by one workgroup in LDS with workgroup-scope arrival
// counters, and the top part (nodes over more than T leaves) in LDS too: its
// maximal subtrees of at most T2 leaves (groups) one workgroup each, then the
// nodes above them (the final top) in a single workgroup.  The stream orders the
// kernels; no device-scope fence is executed.  Every node is still the union of its
// two children in the same operand order as the serial refit, so the boxes are
// bit-identical.
//
// Schedule (int32), for a tree over n leaves:
//   [0] n_sub  [1] n_top  [2] T  [3] top nodes with LDS state per workgroup
//   [4] even lines  [5] n_groups  [6] n_final  [7] T2
//   sub[n]      int4 {subtree root, first leaf, last leaf, up}
//   top_slot[n] internal node -> top slot (-1 below the top part)
//   top_node[n] top slot -> node
//   top_up[n]   top slot -> up link
//   gflag[n]    arrival counters of top slots without LDS state (left at zero;
//               while the schedule is made, the list of top nodes)
//   parity[n]   depth parity of each subtree root (even lines)
//   grp[n]      internal node -> its group (set-up only)
//   goff[n]     group g's top slots are goff[g] .. goff[g+1]-1; the final top's 0 .. n_final-1
//   groot[n]    group g's root node
// An up link is (top slot of the parent) << 1 | (1 if the node is the parent's
// second child), or -1 for the root.
constexpr int kSubMax = 2048;   // leaves of one subtree (LDS of k_refit_sub)
constexpr int kSubBS = 1024;
constexpr int kSubGrid = 2048;
// subtrees of at most 1024 leaves run in a half-size workgroup (77 KB of LDS, so
// two fit on a CU and one's loads overlap the other's walk)
constexpr int kSubHalf = 1024;
constexpr int kSubHalfBS = 512;
// and subtrees of at most 512 leaves in a quarter-size one (39 KB: four per CU)
constexpr int kSubQuarter = 512;
constexpr int kSubQuarterBS = 256;
constexpr int kTopMax = 2048;   // largest refit_top knob (top nodes with LDS state per workgroup)
// The top part's workgroups: 256 threads and LDS state for 512 top nodes (38 KB),
// so they start beside other streams' walks instead of waiting for a whole CU; a
// group holds about kTopGroupLeaves top nodes, and slots past the LDS state count
// arrivals in global memory.
constexpr int kTopBS = 256;
constexpr int kTopCap = 512;
constexpr int kTopOneBS = 1024;   // the whole top part in one workgroup (state for kTopMax nodes)
// a group of the top part spans at most kTopGroupLeaves subtrees' worth of
// leaves (T2 = kTopGroupLeaves * T): about that many top nodes, well within one
// workgroup's LDS state
constexpr int kTopGroupLeaves = 64;
constexpr int kTopGroupGrid = 256;   // workgroups of k_refit_groups (grid-stride over the groups)

struct Sched {
  const int32_t* p;
  int n;
  __device__ int n_sub() const { return p[0]; }
  __device__ int n_top() const { return p[1]; }
  __device__ int top_lds() const { return p[3]; }
  __device__ int4 sub(int i) const { return reinterpret_cast<const int4*>(p + 8)[i]; }
  __device__ int top_slot(int v) const { return p[8 + 4 * (size_t)n + v]; }
  __device__ int top_node(int s) const { return p[8 + 5 * (size_t)n + s]; }
  __device__ int top_up(int s) const { return p[8 + 6 * (size_t)n + s]; }
  __device__ int32_t* gflag(int s) const { return const_cast<int32_t*>(p) + 8 + 7 * (size_t)n + s; }
  // 4-wide lines of even-depth nodes only (single trees: walks start at node 0 and
  // pop only grandchildren, so odd-depth lines are never read)
  __device__ bool even_lines() const { return p[4] != 0; }
  __device__ int sub_parity(int i) const { return p[8 + 8 * (size_t)n + i]; }   // depth of subtree root i, mod 2
  __device__ int n_groups() const { return p[5]; }
  __device__ int n_final() const { return p[6]; }
  __device__ int goff(int g) const { return p[8 + 10 * (size_t)n + g]; }
  __device__ int groot(int g) const { return p[8 + 11 * (size_t)n + g]; }
};

__global__ void k_sched_header(int32_t* sched, int T, int top_lds, int even_lines, int T2) {
  if (threadIdx.x == 0) {
    sched[0] = 0;
    sched[1] = 0;
    sched[2] = T;
    sched[3] = top_lds;
    sched[4] = even_lines;
    sched[5] = 0;
    sched[6] = 0;
    sched[7] = T2;
  }
}

// depth parity of each subtree root (walk to node 0)
__global__ void __launch_bounds__(1024) k_sched_parity(int n, const int32_t* __restrict__ parent,
                                                       int32_t* __restrict__ sched) {
  const size_t N = (size_t)n;
  for (int i = threadIdx.x; i < sched[0]; i += 1024) {
    int node = reinterpret_cast<const int4*>(sched + 8)[i].x, d = 0;
    while (node != 0) {
      node = parent[node];
      ++d;
    }
    sched[8 + 8 * N + i] = d & 1;
  }
}

__device__ __forceinline__ int node_leaves(int j, int n, const int32_t* rl, const int32_t* rr) {
  return j < n - 1 ? rr[j] - rl[j] + 1 : 1;
}

// classify every node: top node (internal, more than T leaves) or subtree root
// (at most T leaves under a top node or at the root)
__global__ void k_sched_mark(int n, const int32_t* __restrict__ parent, const int32_t* __restrict__ rl,
                             const int32_t* __restrict__ rr, int32_t* __restrict__ sched) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= 2 * n - 1) return;
  const int T = sched[2], T2 = sched[7];
  const size_t N = (size_t)n;
  const int sz = node_leaves(j, n, rl, rr);
  if (j < n) sched[8 + 10 * N + j] = 0;          // group sizes, counted next
  if (j < n - 1) {
    int slot = -1;
    if (sz > T) {
      sched[8 + 7 * N + atomicAdd(&sched[1], 1)] = j;   // listed for k_sched_gcount / k_sched_slots
      // its group: the highest ancestor (or itself) over at most T2 leaves; none
      // (the final top) when it has more
      int r = -1;
      if (sz <= T2) {
        r = j;
        for (int p = parent[r]; p >= 0 && node_leaves(p, n, rl, rr) <= T2; p = parent[r]) r = p;
      }
      sched[8 + 9 * N + j] = r;
      if (r == j) {                                // a group root: its group id, kept in top_slot until k_sched_slots
        slot = atomicAdd(&sched[5], 1);
        sched[8 + 11 * N + slot] = j;
      }
    }
    sched[8 + 4 * N + j] = slot;
  }
  const int p = parent[j];
  if (sz <= T && (p < 0 || node_leaves(p, n, rl, rr) > T)) {
    const int e = atomicAdd(&sched[0], 1);
    const int first = j < n - 1 ? rl[j] : j - (n - 1);
    reinterpret_cast<int4*>(sched + 8)[e] = make_int4(j, first, first + sz - 1, p);
  }
}

// top nodes (from the list k_sched_mark made): group ids (the group root's, -1
// for the final top) and group sizes
__global__ void k_sched_gcount(int n, int32_t* __restrict__ sched) {
  const size_t N = (size_t)n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < sched[1]; i += gridDim.x * blockDim.x) {
    const int j = sched[8 + 7 * N + i];
    const int r = sched[8 + 9 * N + j];
    const int g = r < 0 ? -1 : sched[8 + 4 * N + r];
    sched[8 + 9 * N + j] = g;
    atomicAdd(g < 0 ? &sched[6] : &sched[8 + 10 * N + g], 1);
  }
}

// group offsets: the final top's slots first, then each group's (exclusive scan
// of the sizes in one workgroup); the slot fill counters (in top_up, which
// k_sched_link overwrites) zeroed
__global__ void __launch_bounds__(1024) k_sched_gscan(int n, int32_t* __restrict__ sched) {
  __shared__ int s_sum[1024];
  const size_t N = (size_t)n;
  const int ng = sched[5];
  int32_t* goff = sched + 8 + 10 * N;
  int carry = sched[6];
  for (int b = 0; b < ng; b += 1024) {
    const int i = b + (int)threadIdx.x;
    const int v = i < ng ? goff[i] : 0;
    s_sum[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {        // inclusive Hillis-Steele scan
      const int t = (int)threadIdx.x >= o ? s_sum[threadIdx.x - o] : 0;
      __syncthreads();
      s_sum[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < ng) goff[i] = carry + s_sum[threadIdx.x] - v;
    carry += s_sum[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) goff[ng] = carry;
  for (int g = threadIdx.x; g <= ng; g += 1024) sched[8 + 6 * N + g] = 0;
}

// top slots: the final top's and each group's contiguous (order within them
// free); the list's entries become the slots' arrival counters, zeroed (each
// thread reads and clears only its own entry)
__global__ void k_sched_slots(int n, int32_t* __restrict__ sched) {
  const size_t N = (size_t)n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < sched[1]; i += gridDim.x * blockDim.x) {
    const int j = sched[8 + 7 * N + i];
    sched[8 + 7 * N + i] = 0;
    const int g = sched[8 + 9 * N + j], ng = sched[5];
    const int slot = g < 0 ? atomicAdd(&sched[8 + 6 * N + ng], 1)
                           : sched[8 + 10 * N + g] + atomicAdd(&sched[8 + 6 * N + g], 1);
    sched[8 + 4 * N + j] = slot;
    sched[8 + 5 * N + slot] = j;
  }
}

__device__ __forceinline__ int up_link(int j, int p, const int32_t* cb, const int32_t* sched, size_t N) {
  return p < 0 ? -1 : (sched[8 + 4 * N + p] << 1) | (cb[p] == j ? 1 : 0);
}

// up links of subtree roots and top nodes (needs every top slot assigned)
__global__ void k_sched_link(int n, const int32_t* __restrict__ cb, const int32_t* __restrict__ parent,
                             int32_t* __restrict__ sched) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const size_t N = (size_t)n;
  if (i < sched[0]) {
    int4* e = reinterpret_cast<int4*>(sched + 8) + i;
    e->w = up_link(e->x, e->w, cb, sched, N);
  }
  if (i < sched[1]) {
    const int j = sched[8 + 5 * N + i];
    sched[8 + 6 * N + i] = up_link(j, parent[j], cb, sched, N);
  }
}

struct NodeBox {
  float3 lo, hi;
  int num;
};

// lbvh_cell.py:499-515: union of the non-empty children; a node whose children
// are both empty gets the degenerate box 0.25*(min0+max0+min1+max1).  Triangle
// trees have no empty nodes (COUNTED = false).
template <bool COUNTED>
__device__ __forceinline__ NodeBox unite(const NodeBox& a, const NodeBox& b) {
  NodeBox o;
  const int na = COUNTED ? a.num : 1, nb = COUNTED ? b.num : 1;
  if (na > 0 && nb > 0) {
    o.lo = make_float3(fminf(a.lo.x, b.lo.x), fminf(a.lo.y, b.lo.y), fminf(a.lo.z, b.lo.z));
    o.hi = make_float3(fmaxf(a.hi.x, b.hi.x), fmaxf(a.hi.y, b.hi.y), fmaxf(a.hi.z, b.hi.z));
  } else if (na > 0) {
    o.lo = a.lo;
    o.hi = a.hi;
  } else if (nb > 0) {
    o.lo = b.lo;
    o.hi = b.hi;
  } else {
    o.lo = make_float3(0.25f * (((a.lo.x + a.hi.x) + b.lo.x) + b.hi.x), 0.25f * (((a.lo.y + a.hi.y) + b.lo.y) + b.hi.y),
                       0.25f * (((a.lo.z + a.hi.z) + b.lo.z) + b.hi.z));
    o.hi = o.lo;
  }
  o.num = na + nb;
  return o;
}

template <bool COUNTED>
__device__ __forceinline__ NodeBox load_node(int j, const int32_t* num, const float4* nlo, const float4* nhi) {
  NodeBox b;
  b.lo = xyz(nlo[j]);
  b.hi = xyz(nhi[j]);
  b.num = COUNTED ? num[j] : 1;
  return b;
}

// node boxes keep the child ids (internal) or the face / face range (leaf) in .w
template <bool COUNTED>
__device__ __forceinline__ void store_node(int j, const NodeBox& b, int wl, int wr, int32_t* num, float4* nlo,
                                           float4* nhi) {
  nlo[j] = make_float4(b.lo.x, b.lo.y, b.lo.z, __int_as_float(wl));
  nhi[j] = make_float4(b.hi.x, b.hi.y, b.hi.z, __int_as_float(wr));
  if (COUNTED) num[j] = b.num;
}

// 4-wide node array of a triangle tree (float4 quad[8 * internal node], one
// 128-byte line per node): slot 2*side(child) + side(grandchild) of internal node p
// holds the box of p's grandchild as (lo, node index) (hi, face id or -1); when
// p's child is a leaf, its pair of slots holds (that leaf, empty).  Slots 0..3 are
// (left-left, left-right, right-left, right-right): pushed in order and popped
// last-first they visit leaves right-first, the reference's order.  Each internal
// node writes its own line when it is refit.
__device__ __forceinline__ float4 slot_lo(const NodeBox& b, int node) {
  return make_float4(b.lo.x, b.lo.y, b.lo.z, __int_as_float(node));
}
__device__ __forceinline__ float4 slot_hi(const NodeBox& b, int face) {
  return make_float4(b.hi.x, b.hi.y, b.hi.z, __int_as_float(face));
}
__device__ __forceinline__ void empty_slot(float4* d) {
  const float inf = __int_as_float(0x7f800000);
  d[0] = make_float4(inf, inf, inf, __int_as_float(-1));      // overlaps nothing
  d[1] = make_float4(-inf, -inf, -inf, __int_as_float(-1));
}

// where leaf boxes come from
enum LeafKind { kLeafFaceBoxes = 0, kLeafVertices = 1, kLeafCells = 2 };
struct LeafSrc {
  const float4* fmin;       // face boxes (kLeafFaceBoxes, kLeafCells)
  const float4* fmax;
  const int32_t* faces;     // kLeafVertices: boxes from the corners
  const float4* x4;
  float pad;
  const int32_t* sorted;    // triangle: leaf k holds face sorted[k]; cell: faces sorted by cell
  const float* box;         // kLeafCells
  const int32_t* prefix;
  int32_t* leaf_l;
  int32_t* leaf_r;
  float4* quad;             // triangle trees: 4-wide node array (nullable)
  int lines_only;           // triangle trees walked only through their 4-wide lines: of the
                            // internal nodes inside a subtree, store only the subtree root and
                            // its children (what the top part and the top lines read)
};

// Leaf k: its box, written to the node arrays as well.  Cell leaves follow
// assign_leaf_cell_nodes (lbvh_cell.py:601-634): leaf slot k holds row-major cell
// k, its face range from the inclusive prefix, box = cell centre united with its
// faces' boxes.
template <int LEAF>
__device__ __forceinline__ NodeBox leaf_box(int k, int n, const LeafSrc& s, int32_t* num, float4* nlo, float4* nhi) {
  NodeBox b;
  const int node = n - 1 + k;
  if constexpr (LEAF == kLeafCells) {
    const float3 lo = make_float3(s.box[0], s.box[1], s.box[2]);
    const float3 cs = cell_size_of(s.box);
    const int x0 = k / (kGrid * kGrid), y0 = (k % (kGrid * kGrid)) / kGrid, z0 = k % kGrid;
    float3 mn = make_float3(((cs.x * (float)x0) + (0.5f * cs.x)) + lo.x, ((cs.y * (float)y0) + (0.5f * cs.y)) + lo.y,
                            ((cs.z * (float)z0) + (0.5f * cs.z)) + lo.z);
    float3 mx = mn;
    const int l = k ? s.prefix[k - 1] : 0, r = s.prefix[k] - 1;
    for (int j = l; j <= r; ++j) {
      const int f = s.sorted[j];
      const float4 a = s.fmin[f], c = s.fmax[f];
      mn = make_float3(fminf(mn.x, a.x), fminf(mn.y, a.y), fminf(mn.z, a.z));
      mx = make_float3(fmaxf(mx.x, c.x), fmaxf(mx.y, c.y), fmaxf(mx.z, c.z));
    }
    s.leaf_l[k] = l;
    s.leaf_r[k] = r;
    b.lo = mn;
    b.hi = mx;
    b.num = r - l + 1;
    store_node<true>(node, b, l, r, num, nlo, nhi);
  } else {
    const int f = s.sorted[k];
    if constexpr (LEAF == kLeafVertices) {
      face_box(s.faces, s.x4, f, s.pad, b.lo, b.hi);
    } else {
      b.lo = xyz(s.fmin[f]);
      b.hi = xyz(s.fmax[f]);
    }
    b.num = 1;
    store_node<false>(node, b, f, f, num, nlo, nhi);
  }
  return b;
}

// One workgroup per subtree, two leaves per thread (leaves k and k + BS).  Phase 1
// loads, coalesced, the leaves' boxes and parents and the subtree's internal
// nodes' children and parents (all contiguous index ranges) into LDS; phase 2 walks
// up from every leaf, the second child to arrive at a node refitting it, touching
// global memory only to store boxes; phase 3 writes the lines.  Each phase keeps a
// thread's independent chains in flight together: the subtree's phases run one
// after another on its CU, so their latency, not bandwidth, sets the time.
template <int LEAF, int SUBMAX, int BS>
__global__ void __launch_bounds__(BS)
k_refit_sub(int n, const int32_t* __restrict__ sched, LeafSrc src, const int32_t* __restrict__ ca,
            const int32_t* __restrict__ cb, const int32_t* __restrict__ parent, int32_t* num, float4* nlo,
            float4* nhi) {
  static_assert(SUBMAX == 2 * BS, "two leaves per thread");
  constexpr bool COUNTED = LEAF == kLeafCells;
  __shared__ NodeBox s_leaf[SUBMAX];
  __shared__ NodeBox s_int[SUBMAX];
  __shared__ int s_flag[SUBMAX];
  __shared__ int s_ca[SUBMAX], s_cb[SUBMAX], s_par[SUBMAX], s_lface[SUBMAX];
  const Sched S{sched, n};
  const int n_sub = S.n_sub();
  const int tid = (int)threadIdx.x;
  for (int si = blockIdx.x; si < n_sub; si += gridDim.x) {
    const int4 e = S.sub(si);
    const int root = e.x, first = e.y, cnt = e.z - e.y + 1;
    int lpar[2] = {-1, -1};                 // the parents of this thread's leaves
    auto load_links = [&](int k) {          // internal node first+k (those outside the subtree are never read)
      if (k < cnt && first + k < n - 1) {
        s_ca[k] = ca[first + k];
        s_cb[k] = cb[first + k];
        s_par[k] = parent[first + k];
      }
    };
    if constexpr (LEAF == kLeafVertices) {
      // every load issued before any is used (a leaf is three dependent gathers:
      // face id, corner ids, corner positions; its parent and the internal nodes'
      // links are independent of them)
      if (tid < cnt) {
        const int k2 = tid + BS;
        const bool two = k2 < cnt;
        const int f1 = src.sorted[first + tid], f2 = two ? src.sorted[first + k2] : f1;
        lpar[0] = parent[n - 1 + first + tid];
        if (two) lpar[1] = parent[n - 1 + first + k2];
        load_links(tid);
        load_links(k2);
        const int a1 = src.faces[3 * f1], b1 = src.faces[3 * f1 + 1], c1 = src.faces[3 * f1 + 2];
        const int a2 = src.faces[3 * f2], b2 = src.faces[3 * f2 + 1], c2 = src.faces[3 * f2 + 2];
        const float4 pa1 = src.x4[a1], pb1 = src.x4[b1], pc1 = src.x4[c1];
        const float4 pa2 = src.x4[a2], pb2 = src.x4[b2], pc2 = src.x4[c2];
        NodeBox B1, B2;
        corner_box(xyz(pa1), xyz(pb1), xyz(pc1), src.pad, B1.lo, B1.hi);
        corner_box(xyz(pa2), xyz(pb2), xyz(pc2), src.pad, B2.lo, B2.hi);
        B1.num = B2.num = 1;
        s_leaf[tid] = B1;
        s_lface[tid] = f1;
        store_node<false>(n - 1 + first + tid, B1, f1, f1, num, nlo, nhi);
        if (two) {
          s_leaf[k2] = B2;
          s_lface[k2] = f2;
          store_node<false>(n - 1 + first + k2, B2, f2, f2, num, nlo, nhi);
        }
      }
      s_flag[tid] = 0;
      s_flag[tid + BS] = 0;
    } else {
      for (int i = 0; i < 2; ++i) {
        const int k = tid + i * BS;
        s_flag[k] = 0;
        if (k >= cnt) continue;
        lpar[i] = parent[n - 1 + first + k];
        s_leaf[k] = leaf_box<LEAF>(first + k, n, src, num, nlo, nhi);
        if constexpr (!COUNTED) s_lface[k] = src.sorted[first + k];
        load_links(k);
      }
    }
    __syncthreads();
    // phase 2: both leaves' walks advance together
    {
      int node[2], p[2];
      bool act[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        node[i] = n - 1 + first + tid + i * BS;
        act[i] = tid + i * BS < cnt && node[i] != root;
        p[i] = lpar[i];                       // inside the subtree while node != root
      }
      while (act[0] || act[1]) {
        int old[2] = {1, 1};
        __threadfence_block();
#pragma unroll
        for (int i = 0; i < 2; ++i)
          if (act[i]) old[i] = atomicAdd(&s_flag[p[i] - first], 1);
        __threadfence_block();
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          if (!act[i]) continue;
          if (old[i] == 0) {                  // the sibling's walk refits p
            act[i] = false;
            continue;
          }
          const int q = p[i] - first;
          const int l = s_ca[q], r = s_cb[q];
          const NodeBox A = l >= n - 1 ? s_leaf[l - (n - 1) - first] : s_int[l - first];
          const NodeBox B = r >= n - 1 ? s_leaf[r - (n - 1) - first] : s_int[r - first];
          s_int[q] = unite<COUNTED>(A, B);
          node[i] = p[i];
          p[i] = s_par[q];
          act[i] = node[i] != root;
        }
      }
    }
    __syncthreads();
    // the subtree's internal nodes (both children arrived: flag 2) to global memory,
    // coalesced: node boxes one node per thread, quad lines four threads per line
    for (int k = tid; k < cnt; k += BS)
      if (s_flag[k] == 2 && (!src.lines_only || first + k == root || s_par[k] == root))
        store_node<COUNTED>(first + k, s_int[k], s_ca[k], s_cb[k], num, nlo, nhi);
    if constexpr (!COUNTED) {
      if (src.quad && S.even_lines()) {      // odd-depth nodes' lines are never read: flag 2 -> 1
        const int rp = S.sub_parity(si);
        int x[2], d[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int k = tid + i * BS;
          x[i] = (k < cnt && s_flag[k] == 2) ? first + k : root;
          d[i] = rp;
        }
        while (x[0] != root || x[1] != root) {
#pragma unroll
          for (int i = 0; i < 2; ++i)
            if (x[i] != root) {
              x[i] = s_par[x[i] - first];
              ++d[i];
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int k = tid + i * BS;
          if (k < cnt && (d[i] & 1) && s_flag[k] == 2) s_flag[k] = 1;
        }
        __syncthreads();
      }
      if (src.quad) {
        // slot (k, j) of item t = 4k + j: lo and hi of grandchild j of node first+k;
        // a thread's eight items are read together, then stored
        constexpr int IT = 4 * SUBMAX / BS;
        const float inf = __int_as_float(0x7f800000);
        int g[IT];
        bool ok[IT], empty[IT];
#pragma unroll
        for (int i = 0; i < IT; ++i) {
          const int t = tid + i * BS, k = t >> 2, j = t & 3;
          const int kc = t < 4 * cnt ? k : 0;
          const int c = j < 2 ? s_ca[kc] : s_cb[kc];
          ok[i] = t < 4 * cnt && s_flag[kc] == 2;
          empty[i] = c >= n - 1 && (j & 1);   // a leaf child's companion slot
          g[i] = c;
          if (ok[i] && c < n - 1) {
            const int cc = c - first;
            g[i] = (j & 1) ? s_cb[cc] : s_ca[cc];
          }
        }
#pragma unroll
        for (int i = 0; i < IT; ++i) {
          if (!ok[i]) continue;
          const int t = tid + i * BS, k = t >> 2, j = t & 3;
          float4 lo, hi;
          if (empty[i]) {
            lo = make_float4(inf, inf, inf, __int_as_float(-1));
            hi = make_float4(-inf, -inf, -inf, __int_as_float(-1));
          } else {
            const bool leaf = g[i] >= n - 1;
            const NodeBox& B = leaf ? s_leaf[g[i] - (n - 1) - first] : s_int[g[i] - first];
            lo = slot_lo(B, g[i]);
            hi = slot_hi(B, leaf ? s_lface[g[i] - (n - 1) - first] : -1);
          }
          float4* dst = src.quad + 8 * (size_t)(first + k) + 2 * j;
          dst[0] = lo;
          dst[1] = hi;
        }
      }
    }
    __syncthreads();
  }
}

// The top part: one workgroup refits the top slots [s0, s1) (a group, or the
// final top), the first top_lds of them with per-node state in LDS.  Each slot
// with LDS state has two child cells; a finished node puts its box into its
// parent's cell, then arrives at the parent.  The range's children outside it are
// subtree roots and (final top) group roots, whose boxes the earlier kernels
// stored.  Slots beyond the LDS capacity read their children from the node arrays
// and count arrivals in global memory; only this workgroup touches the range's
// slots, so workgroup scope orders both.  A node whose parent is outside the
// range (a group's root) stores its box and ends the walk.
// The 4-wide line of top node p (float4 j of 8: slot j >> 1, low or high half j & 1)
// from the node arrays, once its grandchildren's boxes are final.
__device__ __forceinline__ float4 top_line_entry(int n, int p, int j, const int32_t* __restrict__ ca,
                                                 const int32_t* __restrict__ cb, const float4* nlo,
                                                 const float4* nhi) {
  const int slot = j >> 1, half = j & 1;
  const int c = slot < 2 ? ca[p] : cb[p];
  if (c >= n - 1 && (slot & 1)) {                     // a leaf child's companion slot: empty
    const float inf = __int_as_float(0x7f800000);
    return half ? make_float4(-inf, -inf, -inf, __int_as_float(-1)) : make_float4(inf, inf, inf, __int_as_float(-1));
  }
  const int g = c >= n - 1 ? c : __float_as_int((slot & 1) ? nhi[c].w : nlo[c].w);
  const float4 lo = nlo[g];
  if (half) {
    const float4 hi = nhi[g];
    return make_float4(hi.x, hi.y, hi.z, __int_as_float(g >= n - 1 ? __float_as_int(lo.w) : -1));
  }
  return make_float4(lo.x, lo.y, lo.z, __int_as_float(g));
}

template <bool COUNTED, int CAP, int BS>
__device__ void refit_top_range(const Sched& S, int s0, int s1, const int32_t* __restrict__ ca,
                                const int32_t* __restrict__ cb, int32_t* num, float4* nlo, float4* nhi,
                                float4* quad) {
  __shared__ NodeBox s_cell[CAP][2];
  __shared__ int s_flag[CAP];
  __shared__ int s_up[CAP];
  __shared__ int s_node[CAP], s_l[CAP], s_r[CAP];
  const int cnt = s1 - s0;
  if (cnt <= 0) return;                                 // (the same for the whole workgroup)
  const int cap = S.top_lds() < CAP ? S.top_lds() : CAP;
  const int nl = cnt < cap ? cnt : cap;
  auto inside = [&](int up) { return up >= 0 && (up >> 1) >= s0 && (up >> 1) < s1; };
  for (int i = threadIdx.x; i < nl; i += BS) {
    s_flag[i] = 0;
    s_up[i] = S.top_up(s0 + i);
    const int p = S.top_node(s0 + i);
    s_node[i] = p;
    s_l[i] = ca[p];
    s_r[i] = cb[p];
  }
  // the range's children below it (subtree roots, and group roots for the final
  // top): the children of its nodes that are not its nodes; child b of slot i has
  // up link (s0 + i) << 1 | b.  A slot with LDS state has bit 8 + b of its arrival
  // counter set when child b is in the range (arrivals count in the low byte).
  __syncthreads();
  for (int i = threadIdx.x; i < cnt; i += BS) {
    const int up = i < nl ? s_up[i] : S.top_up(s0 + i);
    if (inside(up) && (up >> 1) - s0 < nl) atomicOr(&s_flag[(up >> 1) - s0], 256 << (up & 1));
  }
  __syncthreads();
  auto kid = [&](int i, int b, int& node) -> bool {
    if (i < nl) {
      node = b ? s_r[i] : s_l[i];
      // (bits 8-9 stay fixed while walks add arrivals to the low byte)
      return !(__hip_atomic_load(&s_flag[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & (256 << b));
    }
    const int c = b ? cb[S.top_node(s0 + i)] : ca[S.top_node(s0 + i)];
    node = c;
    if (c >= S.n - 1) return true;                      // a leaf
    const int ts = S.top_slot(c);
    return ts < s0 || ts >= s1;
  };
  for (int t = threadIdx.x; t < 2 * nl; t += BS) {
    int node;
    if (kid(t >> 1, t & 1, node)) s_cell[t >> 1][t & 1] = load_node<COUNTED>(node, num, nlo, nhi);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 2 * cnt; t += BS) {
    int node;
    if (!kid(t >> 1, t & 1, node)) continue;
    int up = ((s0 + (t >> 1)) << 1) | (t & 1);
    while (inside(up)) {
      const int slot = (up >> 1) - s0;
      const bool lds = slot < nl;
      int prev;
      if (lds) {
        __threadfence_block();
        prev = atomicAdd(&s_flag[slot], 1);
      } else {
        prev = __hip_atomic_fetch_add(S.gflag(s0 + slot), 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      if ((lds ? (prev & 255) : prev) == 0) break;
      int p, l, r;
      NodeBox A, B;
      if (lds) {
        __threadfence_block();
        p = s_node[slot];
        l = s_l[slot];
        r = s_r[slot];
        A = s_cell[slot][0];
        B = s_cell[slot][1];
      } else {
        __hip_atomic_store(S.gflag(s0 + slot), 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        p = S.top_node(s0 + slot);
        l = ca[p];
        r = cb[p];
        A = load_node<COUNTED>(l, num, nlo, nhi);
        B = load_node<COUNTED>(r, num, nlo, nhi);
      }
      const NodeBox P = unite<COUNTED>(A, B);
      up = lds ? s_up[slot] : S.top_up(s0 + slot);
      const bool up_lds = inside(up) && (up >> 1) - s0 < nl;
      // a node whose parent reads the node arrays (or is outside the range)
      // stores now; the others are stored after the walk (a store here would
      // make every level's fence wait for it)
      if (!lds || !up_lds) store_node<COUNTED>(p, P, l, r, num, nlo, nhi);
      if (up_lds) s_cell[(up >> 1) - s0][up & 1] = P;
    }
  }
  __syncthreads();
  // the deferred nodes: both child cells are final, and uniting them again gives
  // the same box
  for (int i = threadIdx.x; i < nl; i += BS) {
    const int up = s_up[i];
    if (inside(up) && (up >> 1) - s0 < nl)
      store_node<COUNTED>(s_node[i], unite<COUNTED>(s_cell[i][0], s_cell[i][1]), s_l[i], s_r[i], num, nlo, nhi);
  }
  // the range's 4-wide lines (triangle trees): every grandchild's box is final
  // now -- in this range (stored above by this workgroup), a subtree root or its
  // child, or a group root (earlier kernels)
  if (!COUNTED && quad) {
    __syncthreads();
    for (int t = threadIdx.x; t < 8 * cnt; t += BS) {
      const int p = S.top_node(s0 + (t >> 3));
      quad[8 * (size_t)p + (t & 7)] = top_line_entry(S.n, p, t & 7, ca, cb, nlo, nhi);
    }
  }
}

// the groups of the top part, one workgroup each (grid-stride over the groups,
// whose count is on the device)
template <bool COUNTED>
__global__ void __launch_bounds__(kTopBS)
k_refit_groups(int n, const int32_t* __restrict__ sched, const int32_t* __restrict__ ca,
               const int32_t* __restrict__ cb, int32_t* num, float4* nlo, float4* nhi, float4* quad) {
  const Sched S{sched, n};
  for (int g = blockIdx.x; g < S.n_groups(); g += gridDim.x) {
    refit_top_range<COUNTED, kTopCap, kTopBS>(S, S.goff(g), S.goff(g + 1), ca, cb, num, nlo, nhi, quad);
    __syncthreads();
  }
}

// the final top (the nodes over more than T2 leaves), in one workgroup
template <bool COUNTED>
__global__ void __launch_bounds__(kTopBS)
k_refit_top(int n, const int32_t* __restrict__ sched, const int32_t* __restrict__ ca,
            const int32_t* __restrict__ cb, int32_t* num, float4* nlo, float4* nhi, float4* quad) {
  const Sched S{sched, n};
  refit_top_range<COUNTED, kTopCap, kTopBS>(S, 0, S.n_final(), ca, cb, num, nlo, nhi, quad);
}

// the whole top part in one workgroup with LDS state for up to kTopMax nodes: the
// groups one after the other, then the final top (trees of at most kTopMax
// subtrees, whose schedule makes the whole top part one group); its lines are
// written by k_top_lines
template <bool COUNTED>
__global__ void __launch_bounds__(kTopOneBS)
k_refit_top_one(int n, const int32_t* __restrict__ sched, const int32_t* __restrict__ ca,
                const int32_t* __restrict__ cb, int32_t* num, float4* nlo, float4* nhi, float4* quad) {
  const Sched S{sched, n};
  for (int g = 0; g < S.n_groups(); ++g) {
    refit_top_range<COUNTED, kTopMax, kTopOneBS>(S, S.goff(g), S.goff(g + 1), ca, cb, num, nlo, nhi, quad);
    __syncthreads();
  }
  refit_top_range<COUNTED, kTopMax, kTopOneBS>(S, 0, S.n_final(), ca, cb, num, nlo, nhi, quad);
}

// The top nodes' 4-wide lines after k_refit_top_one: 8 threads per line (one
// float4 each), grid-stride over the top nodes (their count is on the device);
// one workgroup writing them in turn would wait on their gathers one after another.
__global__ void __launch_bounds__(256)
k_top_lines(int n, const int32_t* __restrict__ sched, const int32_t* __restrict__ ca, const int32_t* __restrict__ cb,
            const float4* __restrict__ nlo, const float4* __restrict__ nhi, float4* __restrict__ quad) {
  const Sched S{sched, n};
  const int n_top = S.n_top();
  for (int t = blockIdx.x * 256 + threadIdx.x; t < 8 * n_top; t += gridDim.x * 256) {
    const int p = S.top_node(t >> 3);
    quad[8 * (size_t)p + (t & 7)] = top_line_entry(n, p, t & 7, ca, cb, nlo, nhi);
  }
}

// One workgroup takes the whole top part when it is likely to have LDS state for
// all its nodes: subtrees of Morton-ordered meshes hold more than half their limit
// on average, so a tree of n leaves has fewer than 2n / refit_sub of them (a
// larger top part is still refit correctly, with arrivals past the LDS state
// counted in global memory, only more slowly).
static inline bool top_in_one(int n) { return 2LL * n <= (long long)kTopMax * knobs().refit_sub; }

size_t refit_sched_ints(int n) { return 8 + 12 * (size_t)(n > 0 ? n : 1) + 8; }

int refit_schedule(int n, const int32_t* cb, const int32_t* parent, const int32_t* rl, const int32_t* rr,
                   int32_t* sched, hipStream_t st, bool even_lines) {
  SLX_LAUNCH("sched", k_sched_header, dim3(1), dim3(64), 0, st, sched, knobs().refit_sub, knobs().refit_top,
             even_lines ? 1 : 0, top_in_one(n) ? n : kTopGroupLeaves * knobs().refit_sub);
  SLX_LAUNCH("sched", k_sched_mark, dim3(blocks_for(2 * n - 1, 256)), dim3(256), 0, st, n, parent, rl, rr, sched);
  const int lg = blocks_for(n, 256) < 256 ? blocks_for(n, 256) : 256;   // grid-stride over the top nodes
  SLX_LAUNCH("sched", k_sched_gcount, dim3(lg), dim3(256), 0, st, n, sched);
  SLX_LAUNCH("sched", k_sched_gscan, dim3(1), dim3(1024), 0, st, n, sched);
  SLX_LAUNCH("sched", k_sched_slots, dim3(lg), dim3(256), 0, st, n, sched);
  SLX_LAUNCH("sched", k_sched_link, dim3(blocks_for(n, 256)), dim3(256), 0, st, n, cb, parent, sched);
  if (even_lines) SLX_LAUNCH("sched", k_sched_parity, dim3(1), dim3(1024), 0, st, n, parent, sched);
  return check_launch("refit_schedule");
}

// after_sub (nullable) is recorded once the subtrees are done: the top part that
// follows is one workgroup, so other streams may take the GPU from there
template <int LEAF>
int refit(const char* name, int n, const int32_t* sched, const LeafSrc& src, const int32_t* ca, const int32_t* cb,
          const int32_t* parent, int32_t* num, float4* nlo, float4* nhi, hipStream_t st,
          hipEvent_t after_sub = nullptr) {
  constexpr bool COUNTED = LEAF == kLeafCells;
  if (knobs().refit_sub <= kSubQuarter) {
    const int grid = n < 4 * kSubGrid ? n : 4 * kSubGrid;
    SLX_LAUNCH(name, (k_refit_sub<LEAF, kSubQuarter, kSubQuarterBS>), dim3(grid), dim3(kSubQuarterBS), 0, st, n,
               sched, src, ca, cb, parent, num, nlo, nhi);
  } else if (knobs().refit_sub <= kSubHalf) {
    const int grid = n < 2 * kSubGrid ? n : 2 * kSubGrid;
    SLX_LAUNCH(name, (k_refit_sub<LEAF, kSubHalf, kSubHalfBS>), dim3(grid), dim3(kSubHalfBS), 0, st, n, sched, src,
               ca, cb, parent, num, nlo, nhi);
  } else {
    const int grid = n < kSubGrid ? n : kSubGrid;
    SLX_LAUNCH(name, (k_refit_sub<LEAF, kSubMax, kSubBS>), dim3(grid), dim3(kSubBS), 0, st, n, sched, src, ca, cb,
               parent, num, nlo, nhi);
  }
  if (after_sub && hipEventRecord(after_sub, st) != hipSuccess) return fail(SLX_EHIP, "refit: event record failed");
  if (top_in_one(n)) {
    SLX_LAUNCH("refit_top", k_refit_top_one<COUNTED>, dim3(1), dim3(kTopOneBS), 0, st, n, sched, ca, cb, num, nlo,
               nhi, nullptr);
    if (!COUNTED && src.quad)
      SLX_LAUNCH("refit_top", k_top_lines, dim3(64), dim3(256), 0, st, n, sched, ca, cb, nlo, nhi, src.quad);
  } else {
    SLX_LAUNCH("refit_groups", k_refit_groups<COUNTED>, dim3(kTopGroupGrid), dim3(kTopBS), 0, st, n, sched, ca, cb,
               num, nlo, nhi, src.quad);
    SLX_LAUNCH("refit_top", k_refit_top<COUNTED>, dim3(1), dim3(kTopBS), 0, st, n, sched, ca, cb, num, nlo, nhi,
               src.quad);
  }
  return check_launch(name);
}

// face -> cell id (assign_face_cell_ids, lbvh_cell.py:576-586) + histogram
__global__ void k_face_cells(int nf, const float4* __restrict__ fmin, const float4* __restrict__ fmax,
                             const float* __restrict__ box, uint32_t* __restrict__ keys,
                             int32_t* __restrict__ ids, int32_t* __restrict__ face_cell,
                             int32_t* __restrict__ hist, int32_t* status) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  float4 a = fmin[f], b = fmax[f];
  float3 p = make_float3(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z));
  int32_t c = flat_cell_id(p, cell_size_of(box), make_float3(box[0], box[1], box[2]));
  if (c < 0 || c >= kCells) {
    if (status) atomicOr(status, kStatusCellRange);
    c = c < 0 ? 0 : kCells - 1;
  }
  keys[f] = (uint32_t)c;
  ids[f] = f;
  face_cell[f] = c;
  atomicAdd(&hist[c], 1);
}

__global__ void k_add_in_place(int n, const int32_t* __restrict__ a, int32_t* __restrict__ b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] += a[i];
}

// ----------------------------------------------------------------- traversal

// Query box of query q: the point q4[q] +- pad, or, with edge queries (qe
// non-null, two vertex ids per query), the box of the edge's end points q4[qe[2q]],
// q4[qe[2q+1]] +- pad (the edge broad phase of XPBD.py:217-232, commented out
// there: min/max of the end points minus/plus padding).
__device__ __forceinline__ void query_box(const float4* __restrict__ q4, const int32_t* __restrict__ qe, int q,
                                          float pad, float3& a, float3& b) {
  const float pp = pad * 1.0f;
  if (qe) {
    const float3 p0 = xyz(q4[qe[2 * q]]), p1 = xyz(q4[qe[2 * q + 1]]);
    a = make_float3(fminf(p0.x, p1.x) - pp, fminf(p0.y, p1.y) - pp, fminf(p0.z, p1.z) - pp);
    b = make_float3(fmaxf(p0.x, p1.x) + pp, fmaxf(p0.y, p1.y) + pp, fmaxf(p0.z, p1.z) + pp);
  } else {
    const float3 p = xyz(q4[q]);
    a = make_float3(p.x - pp, p.y - pp, p.z - pp);
    b = make_float3(p.x + pp, p.y + pp, p.z + pp);
  }
}

// The cell tree's walk split over 16 lanes per query.  The 32^3 cell codes are all
// distinct and differ only in their 15 high interleaved bits, so the Karras tree
// over them is complete: its 16 nodes at depth 4 root subtrees of 2048 cells each.
// Lane s walks the subtree reached from the root by the bits of s (1 = child_b),
// with the sequential walk's rules (skip empty nodes, test boxes, faces of a leaf
// in sorted order); a non-empty node's box lies inside each ancestor's box (the
// union of its non-empty children), so a subtree the sequential walk would reach
// is exactly one whose root overlaps and is not empty.  The sequential walk
// visits child_b's subtree first, so the subtrees come in decreasing s: lane s's
// candidates start after the counts of lanes s+1..15.  Each lane keeps its first
// cap candidates in LDS until the 16 counts are known.
constexpr int kSplit = 16, kSplitQ = 256 / kSplit, kSplitCap = 64, kSplitStack = 16;
__global__ void __launch_bounds__(256)
k_traverse_cell_split(int nq, const float4* __restrict__ q4, float pad, const int32_t* __restrict__ qe,
                      const int32_t* __restrict__ num, const int32_t* __restrict__ sorted,
                      const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                      const float4* __restrict__ fmin, const float4* __restrict__ fmax,
                      const float* __restrict__ box, int cap, int32_t* __restrict__ cand,
                      int32_t* __restrict__ cand_n, int32_t* status, const int32_t* __restrict__ order) {
  __shared__ int32_t s_buf[kSplitCap * 256];      // entry i of lane l at s_buf[i * 256 + l]
  __shared__ int32_t s_stack[kSplitStack * 256];
  const int tid = threadIdx.x, sub = tid & (kSplit - 1);
  const int t = xcd_block() * kSplitQ + (tid >> 4);
  const bool active = t < nq;                     // (16 lanes of one query share it)
  const int q = active ? (order ? order[t] : t) : 0;
  float3 a = make_float3(0.f, 0.f, 0.f), b = a;
  bool live = active;
  if (active) {
    query_box(q4, qe, q, pad, a, b);
    const float3 c = make_float3(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z));
    if (flat_cell_id(c, cell_size_of(box), make_float3(box[0], box[1], box[2])) > kCells - 1) live = false;
  }
  int cnt = 0, flags = 0;
  if (live) {
    int idx = 0;
#pragma unroll
    for (int d = 3; d >= 0; --d) idx = ((sub >> d) & 1) ? __float_as_int(nhi[idx].w) : __float_as_int(nlo[idx].w);
    int32_t* stk = s_stack + tid;
    int sp = 0;
    if (num[idx] > 0) stk[(sp++) * 256] = idx;
    while (sp > 0) {
      const int n = stk[(--sp) * 256];
      const float4 lo = nlo[n], hi = nhi[n];
      if (!overlap(a, b, xyz(lo), xyz(hi))) continue;
      if (n >= kCells - 1) {
        const int l = __float_as_int(lo.w), r = __float_as_int(hi.w);
        for (int j = l; j <= r; ++j) {
          const int f = sorted[j];
          if (overlap(a, b, xyz(fmin[f]), xyz(fmax[f]))) {
            if (cnt < cap) s_buf[cnt * 256 + tid] = f;
            ++cnt;
          }
        }
      } else {
        const int ca = __float_as_int(lo.w), cb = __float_as_int(hi.w);
        if (num[ca] > 0) {
          if (sp < kSplitStack) stk[(sp++) * 256] = ca; else flags |= kStatusStackOverflow;
        }
        if (num[cb] > 0) {
          if (sp < kSplitStack) stk[(sp++) * 256] = cb; else flags |= kStatusStackOverflow;
        }
      }
    }
  }
  // counts of lanes sub..15 (inclusive), within the query's 16 lanes
  int sfx = cnt;
#pragma unroll
  for (int o = 1; o < kSplit; o <<= 1) {
    const int y = __shfl_down(sfx, o, kSplit);
    if (sub + o < kSplit) sfx += y;
  }
  const int total = __shfl(sfx, 0, kSplit), off = sfx - cnt;
  if (!active) return;
  for (int i = 0; i < cnt && i < cap; ++i) {
    const int pos = off + i;
    if (pos < cap) cand[(size_t)pos * nq + q] = s_buf[i * 256 + tid];
  }
  if (sub == 0) cand_n[q] = live ? (total < cap ? total : cap) : 0;
  if (total > cap) flags |= kStatusCandOverflow;
  if (flags && status) atomicOr(status, flags);
}

// traverse_cell_bvh_single, lbvh_cell.py:661-714, for a batch of query points
// p +- pad (or edge boxes, query_box).  CELL: skip queries whose centre cell id exceeds 32767, skip empty
// nodes, leaves list faces [l, r] of sorted; else triangle leaves (one face).
// Children are pushed a then b, so b's subtree is visited first.  Candidates go
// to cand[j * nq + q] (column-major) up to cap; overflow sets a status bit.
template <bool CELL>
__global__ void __launch_bounds__(kTravBS)
k_traverse(int nq, const float4* __restrict__ q4, float pad, const int32_t* __restrict__ qe, int nl,
           const int32_t* __restrict__ num,
           const int32_t* __restrict__ sorted, const float4* __restrict__ nlo, const float4* __restrict__ nhi,
           const float4* __restrict__ fmin, const float4* __restrict__ fmax, const float* __restrict__ box,
           int cap, int32_t* __restrict__ cand, int32_t* __restrict__ cand_n, int32_t* status,
           const int32_t* __restrict__ roots, const int32_t* __restrict__ q_inst, const int32_t* __restrict__ order) {
  __shared__ int32_t s_stack[kStack * kTravBS];
  const int t = xcd_block() * kTravBS + threadIdx.x;
  if (t >= nq) return;
  const int q = order ? order[t] : t;
  int32_t* stk = s_stack + threadIdx.x;          // entry d at stk[d * kTravBS]
  float3 a, b;
  query_box(q4, qe, q, pad, a, b);
  int cnt = 0, flags = 0;
  if (CELL) {
    float3 c = make_float3(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z));
    if (flat_cell_id(c, cell_size_of(box), make_float3(box[0], box[1], box[2])) > kCells - 1) {
      cand_n[q] = 0;
      return;
    }
  }
  if (nl <= 0) {
    cand_n[q] = 0;
    return;
  }
  int sp = 0;
  const int root = roots ? roots[q_inst[q]] : 0;   // batched trees: the query's instance subtree
  if (root < 0) {
    cand_n[q] = 0;
    return;
  }
  stk[0] = root;
  sp = 1;
  while (sp > 0) {
    const int idx = stk[(--sp) * kTravBS];
    if (CELL && num[idx] <= 0) continue;
    const float4 lo = nlo[idx], hi = nhi[idx];
    if (!overlap(a, b, xyz(lo), xyz(hi))) continue;
    if (idx >= nl - 1) {
      const int l = __float_as_int(lo.w), r = __float_as_int(hi.w);
      if (CELL) {
        for (int j = l; j <= r; ++j) {
          const int f = sorted[j];
          if (overlap(a, b, xyz(fmin[f]), xyz(fmax[f]))) {
            if (cnt < cap) cand[(size_t)cnt * nq + q] = f;
            else flags |= kStatusCandOverflow;
            ++cnt;
          }
        }
      } else {
        // triangle leaf: its box is the face box, already tested
        if (cnt < cap) cand[(size_t)cnt * nq + q] = l;
        else flags |= kStatusCandOverflow;
        ++cnt;
      }
    } else {
      const int ca = __float_as_int(lo.w), cb = __float_as_int(hi.w);
      if (!CELL || num[ca] > 0) {
        if (sp < kStack) stk[(sp++) * kTravBS] = ca; else flags |= kStatusStackOverflow;
      }
      if (!CELL || num[cb] > 0) {
        if (sp < kStack) stk[(sp++) * kTravBS] = cb; else flags |= kStatusStackOverflow;
      }
    }
  }
  cand_n[q] = cnt < cap ? cnt : cap;
  if (flags && status) atomicOr(status, flags);
}

}  // namespace slx

using namespace slx;
static inline hipStream_t S(slx_stream_t s) { return (hipStream_t)s; }

extern "C" int slx_face_aabbs(int nf, const int32_t* faces, const float* x4, float padding, float* fmin4,
                              float* fmax4, slx_stream_t stream) {
  SLX_REQUIRE(nf >= 0, "slx_face_aabbs: nf < 0");
  if (nf == 0) return SLX_OK;
  SLX_REQUIRE(faces && x4 && fmin4 && fmax4, "slx_face_aabbs: null pointer");
  SLX_LAUNCH("face_aabbs", k_face_aabbs, dim3(blocks_for(nf, 256)), dim3(256), 0, S(stream), nf, faces,
             (const float4*)x4, padding, (float4*)fmin4, (float4*)fmax4);
  return check_launch("k_face_aabbs");
}

extern "C" size_t slx_mesh_aabb_ws_bytes(void) { return sizeof(float) * 6 * kBoxParts; }

extern "C" int slx_mesh_aabb(int nv, const float* x4, float padding, float* box6, void* ws, size_t ws_bytes,
                             slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0, "slx_mesh_aabb: nv < 0");
  SLX_REQUIRE(box6 && ws && (nv == 0 || x4), "slx_mesh_aabb: null pointer");
  if (ws_bytes < sizeof(float) * 6 * kBoxParts) return fail(SLX_EWS, "slx_mesh_aabb: workspace too small");
  int parts = blocks_for(nv > 0 ? nv : 1, kBoxBS);
  if (parts > kBoxParts) parts = kBoxParts;
  SLX_LAUNCH("box_partial", k_box_partial, dim3(parts), dim3(kBoxBS), 0, S(stream), nv, (const float4*)x4,
             (float*)ws);
  SLX_LAUNCH("box_final", k_box_final, dim3(1), dim3(64), 0, S(stream), parts, (const float*)ws, padding, box6);
  return check_launch("slx_mesh_aabb");
}

extern "C" size_t slx_cell_topology_ws_bytes(void) { return radix_sort_ws_bytes(kCells); }

extern "C" size_t slx_refit_sched_ints(int n_leaves) { return refit_sched_ints(n_leaves); }

extern "C" int slx_set_refit_shape(int sub_leaves, int top_lds) {
  SLX_REQUIRE(sub_leaves >= 1 && sub_leaves <= kSubMax, "slx_set_refit_shape: sub_leaves outside [1, 2048]");
  SLX_REQUIRE(top_lds >= 0 && top_lds <= kTopMax, "slx_set_refit_shape: top_lds outside [0, 2048]");
  knob_defaults().refit_sub = sub_leaves;
  knob_defaults().refit_top = top_lds;
  return SLX_OK;
}

extern "C" int slx_refit_schedule(int n_leaves, const int32_t* child_b, const int32_t* parent,
                                  const int32_t* range_l, const int32_t* range_r, int32_t* sched,
                                  slx_stream_t stream) {
  SLX_REQUIRE(n_leaves >= 1, "slx_refit_schedule: n_leaves < 1");
  SLX_REQUIRE(child_b && parent && range_l && range_r && sched, "slx_refit_schedule: null pointer");
  return refit_schedule(n_leaves, child_b, parent, range_l, range_r, sched, S(stream), false);
}

extern "C" int slx_cell_topology(int32_t* codes, int32_t* sorted_ids, int32_t* child_a, int32_t* child_b,
                                 int32_t* parent, int32_t* range_l, int32_t* range_r, int32_t* sched, void* ws,
                                 size_t ws_bytes, slx_stream_t stream) {
  SLX_REQUIRE(codes && sorted_ids && child_a && child_b && parent && range_l && range_r && sched && ws,
              "slx_cell_topology: null pointer");
  hipStream_t st = S(stream);
  SLX_LAUNCH("cell_morton", k_cell_morton, dim3(blocks_for(kCells, 256)), dim3(256), 0, st,
             make_float3(0.f, 0.f, 0.f), make_float3(1e3f, 1e3f, 1e3f), (uint32_t*)codes, sorted_ids);
  int rc = radix_sort_pairs(kCells, nullptr, (uint32_t*)codes, sorted_ids, 30, ws, ws_bytes, st);
  if (rc) return rc;
  SLX_LAUNCH("karras", k_karras, dim3(blocks_for(kCells, 256)), dim3(256), 0, st, kCells,
             (const uint32_t*)codes, 0, child_a, child_b, parent, range_l, range_r);
  rc = check_launch("slx_cell_topology");
  if (rc) return rc;
  return refit_schedule(kCells, child_b, parent, range_l, range_r, sched, st, false);
}

extern "C" size_t slx_cell_build_ws_bytes(int nf) {
  size_t n = (size_t)(nf > 0 ? nf : 1);
  return 2 * n * 4 + radix_sort_ws_bytes(nf) + scan_ws_bytes(kCells) + 4 * (size_t)kCells * 2 + 1024;
}

extern "C" int slx_cell_build(int nf, const float* fmin4, const float* fmax4, const float* box6,
                              const int32_t* child_a, const int32_t* child_b, const int32_t* parent,
                              const int32_t* sched, int32_t* face_cell, int32_t* prefix, int32_t* sorted_faces,
                              int32_t* num, int32_t* leaf_l, int32_t* leaf_r, float* nlo4, float* nhi4,
                              int32_t* status, void* ws, size_t ws_bytes, slx_stream_t stream) {
  SLX_REQUIRE(nf >= 0, "slx_cell_build: nf < 0");
  SLX_REQUIRE(box6 && child_a && child_b && parent && sched && face_cell && prefix && sorted_faces && num && leaf_l &&
                  leaf_r && nlo4 && nhi4 && ws,
              "slx_cell_build: null pointer");
  SLX_REQUIRE(nf == 0 || (fmin4 && fmax4), "slx_cell_build: null face boxes");
  if (ws_bytes < slx_cell_build_ws_bytes(nf)) return fail(SLX_EWS, "slx_cell_build: workspace too small");
  hipStream_t st = S(stream);
  char* p = (char*)ws;
  size_t n = (size_t)(nf > 0 ? nf : 1);
  uint32_t* keys = (uint32_t*)p; p += 4 * n;
  int32_t* hist = (int32_t*)p; p += 4 * (size_t)kCells;
  p += 4 * (size_t)kCells;
  void* sort_ws = p; size_t sort_bytes = radix_sort_ws_bytes(nf); p += sort_bytes;
  void* scan_ws = p; size_t scan_bytes = scan_ws_bytes(kCells);
  SLX_LAUNCH("fill", k_fill_i32, dim3(blocks_for(kCells, 256)), dim3(256), 0, st, kCells, hist, 0);
  if (nf > 0) {
    SLX_LAUNCH("face_cells", k_face_cells, dim3(blocks_for(nf, 256)), dim3(256), 0, st, nf,
               (const float4*)fmin4, (const float4*)fmax4, box6, keys, sorted_faces, face_cell, hist, status);
    int rc = radix_sort_pairs(nf, nullptr, keys, sorted_faces, 15, sort_ws, sort_bytes, st);
    if (rc) return rc;
  }
  // inclusive prefix = exclusive scan + counts (lbvh_cell.py:644)
  int rc = exclusive_scan(kCells, nullptr, hist, prefix, nullptr, scan_ws, scan_bytes, st);
  if (rc) return rc;
  SLX_LAUNCH("add", k_add_in_place, dim3(blocks_for(kCells, 256)), dim3(256), 0, st, kCells, hist, prefix);
  LeafSrc src{};
  src.fmin = (const float4*)fmin4;
  src.fmax = (const float4*)fmax4;
  src.sorted = sorted_faces;
  src.box = box6;
  src.prefix = prefix;
  src.leaf_l = leaf_l;
  src.leaf_r = leaf_r;
  return refit<kLeafCells>("cell_refit", kCells, sched, src, child_a, child_b, parent, num, (float4*)nlo4,
                           (float4*)nhi4, st);
}

extern "C" size_t slx_tri_build_ws_bytes(int nf) {
  size_t n = (size_t)(nf > 0 ? nf : 1);
  return 2 * n * 4 + radix_sort_ws_bytes(nf) + 1024;
}

extern "C" int slx_tri_refit(int nf, const int32_t* faces, const float* x4, float padding,
                             const int32_t* sorted_faces, const int32_t* child_a, const int32_t* child_b,
                             const int32_t* parent, const int32_t* sched, float* nlo4, float* nhi4, float* quad4,
                             slx_stream_t stream) {
  return tri_refit_impl(nf, faces, x4, padding, sorted_faces, child_a, child_b, parent, sched, nlo4, nhi4, quad4,
                        S(stream), nullptr);
}

int slx::tri_refit_impl(int nf, const int32_t* faces, const float* x4, float padding,
                        const int32_t* sorted_faces, const int32_t* child_a, const int32_t* child_b, const int32_t* parent, const int32_t* sched,
                   float* nlo4, float* nhi4, float* quad4, hipStream_t st, hipEvent_t after_sub,
                   bool lines_only) {
  SLX_REQUIRE(nf >= 0, "slx_tri_refit: nf < 0");
  if (nf == 0) return SLX_OK;
  SLX_REQUIRE(faces && x4 && sorted_faces && child_a && child_b && parent && sched && nlo4 && nhi4,
              "slx_tri_refit: null pointer");
  LeafSrc src{};
  src.faces = faces;
  src.x4 = (const float4*)x4;
  src.pad = padding;
  src.sorted = sorted_faces;
  src.quad = (float4*)quad4;
  src.lines_only = lines_only && quad4 ? 1 : 0;
  return refit<kLeafVertices>("tri_refit", nf, sched, src, child_a, child_b, parent, nullptr, (float4*)nlo4,
                              (float4*)nhi4, st, after_sub);
}

// The leaves' boxes alone (leaf k: face sorted[k], its corners' box grown by pad),
// computed and stored exactly as the subtree refit's first phase does.
__global__ void __launch_bounds__(256)
k_leaf_boxes(int nf, const int32_t* __restrict__ faces, const float4* __restrict__ x4, float pad,
             const int32_t* __restrict__ sorted, float4* __restrict__ nlo, float4* __restrict__ nhi) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  const int f = sorted[k];
  NodeBox B;
  corner_box(xyz(x4[faces[3 * f]]), xyz(x4[faces[3 * f + 1]]), xyz(x4[faces[3 * f + 2]]), pad, B.lo, B.hi);
  B.num = 1;
  store_node<false>(nf - 1 + k, B, f, f, nullptr, nlo, nhi);
}

int slx::tri_leaf_boxes_impl(int nf, const int32_t* faces, const float* x4, float padding,
                             const int32_t* sorted_faces, float* nlo4, float* nhi4, hipStream_t st) {
  SLX_REQUIRE(nf >= 0, "tri_leaf_boxes: nf < 0");
  if (nf == 0) return SLX_OK;
  SLX_REQUIRE(faces && x4 && sorted_faces && nlo4 && nhi4, "tri_leaf_boxes: null pointer");
  SLX_LAUNCH("tri_leaf_boxes", k_leaf_boxes, dim3(blocks_for(nf, 256)), dim3(256), 0, st, nf, faces,
             (const float4*)x4, padding, sorted_faces, (float4*)nlo4, (float4*)nhi4);
  return check_launch("tri_leaf_boxes");
}

extern "C" int slx_tri_build(int nf, const float* fmin4, const float* fmax4, const float* box6, int32_t* codes,
                             int32_t* sorted_faces, int32_t* child_a, int32_t* child_b, int32_t* parent,
                             int32_t* range_l, int32_t* range_r, int32_t* sched, float* nlo4, float* nhi4,
                             float* quad4, void* ws, size_t ws_bytes, slx_stream_t stream) {
  SLX_REQUIRE(nf >= 0, "slx_tri_build: nf < 0");
  if (nf == 0) return SLX_OK;
  SLX_REQUIRE(fmin4 && fmax4 && box6 && codes && sorted_faces && child_a && child_b && parent && range_l && range_r &&
                  sched && nlo4 && nhi4 && ws,
              "slx_tri_build: null pointer");
  if (ws_bytes < slx_tri_build_ws_bytes(nf)) return fail(SLX_EWS, "slx_tri_build: workspace too small");
  hipStream_t st = S(stream);
  SLX_LAUNCH("face_morton", k_face_morton, dim3(blocks_for(nf, 256)), dim3(256), 0, st, nf,
             (const float4*)fmin4, (const float4*)fmax4, box6, (uint32_t*)codes, sorted_faces);
  int rc = radix_sort_pairs(nf, nullptr, (uint32_t*)codes, sorted_faces, 30, (char*)ws + 8 * (size_t)nf + 512,
                            ws_bytes - 8 * (size_t)nf - 512, st);
  if (rc) return rc;
  SLX_LAUNCH("karras", k_karras, dim3(blocks_for(nf, 256)), dim3(256), 0, st, nf, (const uint32_t*)codes, 1,
             child_a, child_b, parent, range_l, range_r);
  rc = check_launch("k_karras");
  if (rc) return rc;
  rc = refit_schedule(nf, child_b, parent, range_l, range_r, sched, st, true);
  if (rc) return rc;
  LeafSrc src{};
  src.fmin = (const float4*)fmin4;
  src.fmax = (const float4*)fmax4;
  src.sorted = sorted_faces;
  src.quad = (float4*)quad4;
  return refit<kLeafFaceBoxes>("tri_refit", nf, sched, src, child_a, child_b, parent, nullptr, (float4*)nlo4,
                               (float4*)nhi4, st);
}

namespace slx {

// Triangle-tree traversal (traverse_cell_bvh_single's stack walk,
// lbvh_cell.py:661-714, on the per-triangle tree): a popped node's two children
// are tested together and only overlapping ones are pushed.  In a Karras tree the
// children of an internal node are split and split+1 (internal, or the leaves
// n-1+split, n-1+split+1), so their boxes are neighbours in memory and one stack
// word names both: the left child index, with bit 31 set when the right child is
// a leaf.  A left leaf that overlaps is pushed as an emit entry (face | bit 30) so
// that it is reported after the right subtree, exactly as in the reference's order
// (push left, push right, pop right first).  Candidate lists therefore match the
// push-everything walk entry for entry; the stack only ever holds overlapping
// nodes, so it is never deeper than that walk's.
constexpr uint32_t kRightLeaf = 1u << 31, kEmit = 1u << 30, kLow30 = kEmit - 1;

__global__ void __launch_bounds__(kTravBS)
k_traverse_tri(int nq, const float4* __restrict__ q4, float pad, const int32_t* __restrict__ qe, int nl,
               const float4* __restrict__ nlo,
               const float4* __restrict__ nhi, int cap, int32_t* __restrict__ cand, int32_t* __restrict__ cand_n,
               int32_t* status, const int32_t* __restrict__ roots, const int32_t* __restrict__ q_inst,
               const int32_t* __restrict__ order, const int32_t* __restrict__ n_active) {
  __shared__ uint32_t s_stack[kStack * kTravBS];
  const int t = (n_active ? (int)blockIdx.x : xcd_block()) * kTravBS + threadIdx.x;
  if (t >= nq || (n_active && t >= *n_active)) return;
  const int q = order ? order[t] : t;     // spatially coherent queries share nodes within a wavefront
  uint32_t* stk = s_stack + threadIdx.x;          // entry d at stk[d * kTravBS]
  float3 a, b;
  query_box(q4, qe, q, pad, a, b);
  int cnt = 0, flags = 0;
  const int root = roots ? roots[q_inst[q]] : 0;
  if (nl <= 0 || root < 0) {
    cand_n[q] = 0;
    return;
  }
  auto emit = [&](int f) {
    if (cnt < cap) cand[(size_t)cnt * nq + q] = f;
    else flags |= kStatusCandOverflow;
    ++cnt;
  };
  const int leaf0 = nl - 1;
  auto expand_word = [&](int l, int r) { return (uint32_t)l | (r >= leaf0 ? kRightLeaf : 0u); };
  int sp = 0;
  {
    const float4 lo = nlo[root], hi = nhi[root];
    if (overlap(a, b, xyz(lo), xyz(hi))) {
      if (root >= leaf0) emit(__float_as_int(lo.w));
      else stk[(sp++) * kTravBS] = expand_word(__float_as_int(lo.w), __float_as_int(hi.w));
    }
  }
  while (sp > 0) {
    const uint32_t e = stk[(--sp) * kTravBS];
    if (e & kEmit) {
      emit((int)(e & kLow30));
      continue;
    }
    const int l = (int)(e & kLow30);
    const int split = l >= leaf0 ? l - leaf0 : l;
    const int r = (e & kRightLeaf) ? leaf0 + split + 1 : split + 1;
    const float4 l0 = nlo[l], l1 = nhi[l], r0 = nlo[r], r1 = nhi[r];
    if (overlap(a, b, xyz(l0), xyz(l1))) {
      const uint32_t w = l >= leaf0 ? ((uint32_t)__float_as_int(l0.w) | kEmit)
                                    : expand_word(__float_as_int(l0.w), __float_as_int(l1.w));
      if (sp < kStack) stk[(sp++) * kTravBS] = w;
      else flags |= kStatusStackOverflow;
    }
    if (overlap(a, b, xyz(r0), xyz(r1))) {
      if (r >= leaf0) {
        emit(__float_as_int(r0.w));
      } else if (sp < kStack) {
        stk[(sp++) * kTravBS] = expand_word(__float_as_int(r0.w), __float_as_int(r1.w));
      } else {
        flags |= kStatusStackOverflow;
      }
    }
  }
  cand_n[q] = cnt < cap ? cnt : cap;
  if (flags && status) atomicOr(status, flags);
}

// Wavefront-coherent variant of k_traverse_tri: the 64 lanes of a wavefront (64
// queries, taken in a spatially coherent order) walk one shared depth-first
// traversal.  A child is pushed when any lane's box overlaps it, node boxes are
// wave-uniform scalar loads, and the stack is wave-uniform too: entry d lives in
// lane d's register (64 entries, as the reference's stack), read back with
// readlane, so no LDS is used.  Each lane emits a leaf only when its own box
// overlaps it; every ancestor of such a leaf overlaps that box too, so the leaves a
// lane emits, and their order, are exactly those of its own walk (the shared walk
// visits a superset of its nodes in the same right-first order).
struct CacheOut {
  float pad_walk;          // padding of the enlarged box
  int cap;                 // cache capacity per query
  int32_t* leaves;         // [cap][nq] by slot
  int32_t* count;          // [nq] by slot
  float4* q0;              // [nq] by slot (w: the query's own-face count, see OwnFaces)
  int32_t* redo;           // [1 + nq]
  const int4* own;         // [nq][2] by query (nullable): faces each query leaves out of its lists
  const uint32_t* stale;   // (per-lane walks, nullable) displacement bound of internal boxes older than the leaves'
};

// Own faces of a cloth vertex queried against its own mesh's tree: up to eight
// face ids (-1 padded), the faces incident to the vertex, which the narrow phase
// skips (the is_in_face test, XPBD.py:258).  A walk that has them counts such a
// face in the list's length (so a list overflows exactly when the full list
// does, and the faces it keeps are those the full list keeps) but stores only
// the others, in the same order: the contacts, their order and so every sum are
// unchanged, while the cloth's lists lose most of their entries (six of the
// eight a draped vertex has).  cand_n then counts the stored entries.  The cache
// build leaves them out of the stored leaves too and writes their number into
// q0[t].w, which bounds how many of them a later filter could have found.
struct OwnFaces {
  int4 a = make_int4(-1, -1, -1, -1), b = make_int4(-1, -1, -1, -1);
  __device__ __forceinline__ void load(const int4* own, int q) {
    if (own) {
      a = own[2 * (size_t)q];
      b = own[2 * (size_t)q + 1];
    }
  }
  __device__ __forceinline__ bool has(int f) const {
    return ((f == a.x || f == a.y) || (f == a.z || f == a.w)) || ((f == b.x || f == b.y) || (f == b.z || f == b.w));
  }
  __device__ __forceinline__ int count() const {
    return ((a.x >= 0) + (a.y >= 0) + (a.z >= 0) + (a.w >= 0)) + ((b.x >= 0) + (b.y >= 0) + (b.z >= 0) + (b.w >= 0));
  }
};

// Stored leaves go out four at a time: a lane collects them in an int4 and writes
// it to row (i / 4) of the int4 array [(cap + 3) / 4][nq] when it is full (and
// the partial last one at the end), so a store covers 16 bytes, not 4.
struct LeafPack {
  int4 v = make_int4(0, 0, 0, 0);
  __device__ __forceinline__ void put(int i, int node, int cap, int nq, int t, int32_t* leaves) {
    if (i >= cap) return;
    const int s = i & 3;
    v.x = s == 0 ? node : v.x;
    v.y = s == 1 ? node : v.y;
    v.z = s == 2 ? node : v.z;
    v.w = s == 3 ? node : v.w;
    if (s == 3) reinterpret_cast<int4*>(leaves)[(size_t)(i >> 2) * nq + t] = v;
  }
  __device__ __forceinline__ void finish(int n, int cap, int nq, int t, int32_t* leaves) {
    const int m = n < cap ? n : cap;
    if (m & 3) reinterpret_cast<int4*>(leaves)[(size_t)(m >> 2) * nq + t] = v;
  }
};

constexpr uint32_t kExactBit = 1u << 31;   // quad walk, cache build: with kEmit, the leaf overlaps the exact box
constexpr uint32_t kCacheNode = 1u << 31;  // quad walk, cache build: without kEmit, a leaf node for the cache

// BUILD and n_active as for k_traverse_quad (the cache arrays in CacheOut).
template <bool BUILD>
__global__ void __launch_bounds__(256)
k_traverse_tri_wave(int nq, const float4* __restrict__ q4, float pad, const int32_t* __restrict__ qe, int nl,
                    const float4* __restrict__ nlo,
                    const float4* __restrict__ nhi, int cap, int32_t* __restrict__ cand,
                    int32_t* __restrict__ cand_n, int32_t* status, const int32_t* __restrict__ order,
                    const int32_t* __restrict__ n_active, CacheOut co) {
  // a grid-stride loop (launches sized by n_active get a capped grid); the step is
  // per workgroup, so every lane of a wavefront takes the same iterations
  const int limit = n_active ? min(nq, *n_active) : nq;
  for (int base = (n_active ? (int)blockIdx.x : xcd_block()) * 256; base < limit; base += gridDim.x * 256)
  [&](const int t) {
  const int lane = threadIdx.x & 63;
  const bool active = t < limit;
  const int q = active ? (order ? order[t] : t) : 0;
  float3 a, b, aw, bw;
  if (active) {
    query_box(q4, qe, q, pad, a, b);
    aw = a;
    bw = b;
    if (BUILD) {                             // point queries only
      const float3 p = xyz(q4[q]);
      const float pw = co.pad_walk;
      aw = make_float3(p.x - pw, p.y - pw, p.z - pw);
      bw = make_float3(p.x + pw, p.y + pw, p.z + pw);
      co.q0[t] = make_float4(p.x, p.y, p.z, 0.f);
    }
  } else {                                   // an empty box: overlaps nothing
    const float inf = __int_as_float(0x7f800000);
    a = aw = make_float3(inf, inf, inf);
    b = bw = make_float3(-inf, -inf, -inf);
  }
  int cnt = 0, ccnt = 0, flags = 0;
  auto emit = [&](int f) {
    if (cnt < cap) cand[(size_t)cnt * nq + q] = f;
    else flags |= kStatusCandOverflow;
    ++cnt;
  };
  // a leaf reached by the walk: to the cache (BUILD) and, if it overlaps the exact box, to the list
  LeafPack pack;
  auto leaf = [&](int node, const float4& lo, const float4& hi) {
    if (BUILD) {
      pack.put(ccnt, node, co.cap, nq, t, co.leaves);
      ++ccnt;
      if (overlap(a, b, xyz(lo), xyz(hi))) emit(__float_as_int(lo.w));
    } else {
      emit(__float_as_int(lo.w));
    }
  };
  const int leaf0 = nl - 1;
  auto expand_word = [&](int l, int r) { return (uint32_t)l | (r >= leaf0 ? kRightLeaf : 0u); };
  uint32_t sreg = 0;                         // this lane's stack slot
  int sp = 0;                                // wave-uniform depth
  auto push = [&](uint32_t v) {
    if (sp < kWave) {
      if (lane == sp) sreg = v;
      ++sp;
    } else {
      flags |= kStatusStackOverflow;
    }
  };
  if (nl > 0) {
    const float4 lo = nlo[0], hi = nhi[0];
    const bool ov = overlap(aw, bw, xyz(lo), xyz(hi));
    if (__ballot(ov)) {
      if (leaf0 == 0) {
        if (ov) leaf(0, lo, hi);
      } else {
        push(expand_word(__float_as_int(lo.w), __float_as_int(hi.w)));
      }
    }
  }
  while (__builtin_amdgcn_readfirstlane(sp) > 0) {
    sp = __builtin_amdgcn_readfirstlane(sp) - 1;
    const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)sreg, sp);
    if (e & kEmit) {                         // a left leaf, reported after its right sibling's subtree
      const int L = (int)(e & kLow30);
      const float4 lo = nlo[L], hi = nhi[L];
      if (overlap(aw, bw, xyz(lo), xyz(hi))) leaf(L, lo, hi);
      continue;
    }
    const int l = (int)(e & kLow30);
    const int split = l >= leaf0 ? l - leaf0 : l;
    const int r = (e & kRightLeaf) ? leaf0 + split + 1 : split + 1;
    const float4 l0 = nlo[l], l1 = nhi[l], r0 = nlo[r], r1 = nhi[r];
    const bool ovl = overlap(aw, bw, xyz(l0), xyz(l1));
    const bool ovr = overlap(aw, bw, xyz(r0), xyz(r1));
    if (__ballot(ovl))
      push(l >= leaf0 ? ((uint32_t)l | kEmit) : expand_word(__float_as_int(l0.w), __float_as_int(l1.w)));
    if (__ballot(ovr)) {
      if (r >= leaf0) {
        if (ovr) leaf(r, r0, r1);
      } else {
        push(expand_word(__float_as_int(r0.w), __float_as_int(r1.w)));
      }
    }
  }
  if (!active) return;
  if (BUILD) {
    const bool over = flags & kStatusStackOverflow;
    pack.finish(ccnt, co.cap, nq, t, co.leaves);
    co.count[t] = over ? INT_MAX : ccnt;
    if (over) {
      co.redo[1 + atomicAdd(&co.redo[0], 1)] = q;
      return;
    }
    flags &= ~kStatusStackOverflow;
  }
  cand_n[q] = cnt < cap ? cnt : cap;
  if (flags && status) atomicOr(status, flags);
  }(base + (int)threadIdx.x);
}

}  // namespace slx

namespace slx {

// Wavefront walk over the 4-wide node array, for fine trees whose queries come in
// a coherent order (neighbouring vertices in a wavefront): the 64 lanes share one
// depth-first walk over the union of their boxes, with scalar line loads and the
// stack one entry per lane register.  A popped line's four grandchild slots are
// tested by every lane against its own box.  Slots are visited in the order 3, 2,
// 1, 0 (the reference's right-first order, as in k_traverse_quad): the leaf slots
// that come before the line's last internal slot in that order are reported at
// once, the internal slots and the leaves after them are pushed in slot order (a
// pushed leaf, tagged kCacheNode, is reported when popped, its box reloaded).
// Each lane reports only the leaves its own box overlaps, in the shared walk's
// order, which restricted to them is its own walk's order: every lane's list is
// the one the per-lane walk gives.  BUILD, co and n_active as for k_traverse_quad;
// a stack overflow (shared) sends every active lane of the wavefront to co.redo.
template <bool BUILD>
__global__ void __launch_bounds__(256)
k_traverse_quad_wave(int nq, const float4* __restrict__ q4, float pad, const int32_t* __restrict__ qe, int nl,
                     const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                     const float4* __restrict__ quad, int cap, int32_t* __restrict__ cand,
                     int32_t* __restrict__ cand_n, int32_t* status, const int32_t* __restrict__ order,
                     const int32_t* __restrict__ n_active, CacheOut co) {
  const int limit = n_active ? min(nq, *n_active) : nq;
  for (int base = (n_active ? (int)blockIdx.x : xcd_block()) * 256; base < limit; base += gridDim.x * 256)
  [&](const int t) {
  const int lane = threadIdx.x & 63;
  const bool active = t < limit;
  const int q = active ? (order ? order[t] : t) : 0;
  float3 a, b, aw, bw;
  OwnFaces own;
  if (active) {
    own.load(co.own, q);
    query_box(q4, qe, q, pad, a, b);
    aw = a;
    bw = b;
    if (BUILD) {                             // point queries only
      const float3 p = xyz(q4[q]);
      const float pw = co.pad_walk;
      aw = make_float3(p.x - pw, p.y - pw, p.z - pw);
      bw = make_float3(p.x + pw, p.y + pw, p.z + pw);
      co.q0[t] = make_float4(p.x, p.y, p.z, __int_as_float(own.count()));
    }
  } else {                                   // an empty box: overlaps nothing
    const float inf = __int_as_float(0x7f800000);
    a = aw = make_float3(inf, inf, inf);
    b = bw = make_float3(-inf, -inf, -inf);
  }
  int all = 0, kept = 0, ccnt = 0, flags = 0;
  auto emit = [&](int f) {                   // own faces counted, not stored (OwnFaces)
    if (all >= cap) flags |= kStatusCandOverflow;
    else if (!own.has(f)) cand[(size_t)(kept++) * nq + q] = f;
    ++all;
  };
  // a leaf (node id, face id, box) whose box this lane's walk box overlaps: to the
  // cache (BUILD) and, if it overlaps the exact box, to the list.  (A line slot
  // holds the node id in lo.w and the face in hi.w; the binary node arrays hold
  // a leaf's face in lo.w.)
  LeafPack pack;
  auto leaf = [&](int node, int face, const float4& lo, const float4& hi) {
    if (BUILD) {
      if (!own.has(face)) {
        pack.put(ccnt, node, co.cap, nq, t, co.leaves);
        ++ccnt;
      }
      if (overlap(a, b, xyz(lo), xyz(hi))) emit(face);
    } else {
      emit(face);
    }
  };
  uint32_t sreg = 0;                         // this lane's stack slot
  int sp = 0;                                // wave-uniform depth
  auto push = [&](uint32_t v) {
    if (sp < kWave) {
      if (lane == sp) sreg = v;
      ++sp;
    } else {
      flags |= kStatusStackOverflow;
    }
  };
  if (nl > 0) {
    const float4 lo = nlo[0], hi = nhi[0];
    const bool ov = overlap(aw, bw, xyz(lo), xyz(hi));
    if (__ballot(ov)) {
      if (nl == 1) {
        if (ov) leaf(0, __float_as_int(lo.w), lo, hi);
      } else {
        push(0u);
      }
    }
  }
  while (__builtin_amdgcn_readfirstlane(sp) > 0) {
    sp = __builtin_amdgcn_readfirstlane(sp) - 1;
    const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)sreg, sp);
    if (e & kCacheNode) {                    // a leaf pushed below an internal slot
      const int L = (int)(e & kLow30);
      const float4 lo = nlo[L], hi = nhi[L];
      if (overlap(aw, bw, xyz(lo), xyz(hi))) leaf(L, __float_as_int(lo.w), lo, hi);
      continue;
    }
    const float4* qn = quad + 8 * (size_t)e;
    float4 lo[4], hi[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      lo[k] = qn[2 * k];
      hi[k] = qn[2 * k + 1];
    }
    bool ov[4], any[4];
    int face[4], node[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      ov[k] = overlap(aw, bw, xyz(lo[k]), xyz(hi[k]));
      any[k] = __ballot(ov[k]) != 0ull;
      face[k] = __float_as_int(hi[k].w);
      node[k] = __float_as_int(lo[k].w);
    }
    // the last internal slot some lane overlaps, in visiting order 3, 2, 1, 0
    int kstar = -1;
#pragma unroll
    for (int k = 3; k >= 0; --k)
      if (kstar < 0 && any[k] && face[k] < 0) kstar = k;
#pragma unroll
    for (int k = 3; k >= 0; --k)
      if (k > kstar && ov[k] && face[k] >= 0) leaf(node[k], face[k], lo[k], hi[k]);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k <= kstar && any[k]) push(face[k] >= 0 ? ((uint32_t)node[k] | kCacheNode) : (uint32_t)node[k]);
    if (co.redo && (flags & kStatusStackOverflow)) break;
  }
  if (!active) return;
  if (co.redo && (flags & kStatusStackOverflow)) {     // walked again by the second tier
    if (BUILD) co.count[t] = INT_MAX;
    co.redo[1 + atomicAdd(&co.redo[0], 1)] = q;
    return;
  }
  if (BUILD) {
    pack.finish(ccnt, co.cap, nq, t, co.leaves);
    co.count[t] = ccnt;
  }
  cand_n[q] = kept;
  if (flags && status) atomicOr(status, flags);
  }(base + (int)threadIdx.x);
}

// Triangle-tree walk over the 4-wide node array: a popped node's four grandchild
// slots (one 128-byte line) are tested together, so one dependent load covers two
// tree levels.  Overlapping slots are pushed in slot order (a leaf as an emit
// entry), so they pop right-first: the leaves come out in the reference's order
// (the right subtree before the left, i.e. in decreasing leaf position), and the
// set of emitted leaves is the same, since a grandchild overlaps the query box only
// if its parent does (a parent's box is the union of its children's).
//
// BUILD (candidate cache, see k_cand_filter): the walk uses the box enlarged by
// the cache's margin; every leaf it reaches goes to the cache (node ids, first
// cache cap, indexed by processing slot t so a wavefront's accesses coalesce) and
// those that overlap the exact box also to the exact list.  A leaf slot pushes
// two words, its node id (tagged kCacheNode) and then its face (tagged kEmit, with
// kExactBit when it overlaps the exact box), so both come off the stack together
// without a load.  A stack overflow marks the cache entry incomplete and sends
// the query to the redo list.
// n_active (nullable, device): only the first *n_active entries of order[] are
// walked (a list of queries built on the device).

//
// Stack depth S: with co.redo set (a first-tier walk), a query whose stack
// overflows is appended to co.redo instead (its lists are then written by the
// second walk); otherwise the overflow sets the status bit.
template <bool BUILD, int S>
__global__ void __launch_bounds__(kTravBS)
k_traverse_quad(int nq, const float4* __restrict__ q4, float pad, const int32_t* __restrict__ qe, int nl,
                const float4* __restrict__ nlo,
                const float4* __restrict__ nhi, const float4* __restrict__ quad, int cap, int32_t* __restrict__ cand,
                int32_t* __restrict__ cand_n, int32_t* status, const int32_t* __restrict__ roots,
                const int32_t* __restrict__ q_inst, const int32_t* __restrict__ order,
                const int32_t* __restrict__ n_active, CacheOut co) {
  __shared__ uint32_t s_stack[S * kTravBS];
  // a grid-stride loop (launches sized by n_active get a capped grid)
  const int limit = n_active ? min(nq, *n_active) : nq;
  for (int t0 = (n_active ? (int)blockIdx.x : xcd_block()) * kTravBS + threadIdx.x; t0 < limit;
       t0 += gridDim.x * kTravBS) [&](const int t) {
  const int q = order ? order[t] : t;
  uint32_t* stk = s_stack + threadIdx.x;          // entry d at stk[d * kTravBS]
  float3 a, b;
  query_box(q4, qe, q, pad, a, b);
  float3 aw = a, bw = b;                          // the walk's box
  if (!BUILD && co.stale) {
    // internal boxes from positions up to D older: the walk's box grows by D (and
    // a margin for the boxes' rounding), and leaves are tested by their current boxes
    const float mag = fmaxf(fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(b.x))),
                            fmaxf(fabsf(b.y), fabsf(b.z)));
    const float e = __uint_as_float(*co.stale) + ldexpf(mag, -20);
    aw = make_float3(a.x - e, a.y - e, a.z - e);
    bw = make_float3(b.x + e, b.y + e, b.z + e);
  }
  OwnFaces own;
  own.load(co.own, q);
  if (BUILD) {                                    // point queries only
    const float3 p = xyz(q4[q]);
    const float pw = co.pad_walk;
    aw = make_float3(p.x - pw, p.y - pw, p.z - pw);
    bw = make_float3(p.x + pw, p.y + pw, p.z + pw);
    co.q0[t] = make_float4(p.x, p.y, p.z, __int_as_float(own.count()));
  }
  int all = 0, kept = 0, ccnt = 0, flags = 0;
  const int root = roots ? roots[q_inst[q]] : 0;
  if (nl <= 0 || root < 0) {
    cand_n[q] = 0;
    if (BUILD) co.count[t] = 0;
    return;
  }
  // the full list's first cap entries, own faces counted but not stored
  auto emit = [&](int f) {
    if (all >= cap) flags |= kStatusCandOverflow;
    else if (!own.has(f)) cand[(size_t)(kept++) * nq + q] = f;
    ++all;
  };
  LeafPack pack;
  auto keep = [&](int node) {
    pack.put(ccnt, node, co.cap, nq, t, co.leaves);
    ++ccnt;
  };
  int sp = 0;
  {
    const float4 lo = nlo[root], hi = nhi[root];
    if (overlap(aw, bw, xyz(lo), xyz(hi))) {
      if (root >= nl - 1) {
        if (BUILD && !own.has(__float_as_int(lo.w))) keep(root);
        if ((!BUILD && !co.stale) || overlap(a, b, xyz(lo), xyz(hi))) emit(__float_as_int(lo.w));
      } else {
        stk[(sp++) * kTravBS] = (uint32_t)root;
      }
    }
  }
  while (sp > 0) {
    const uint32_t e = stk[(--sp) * kTravBS];
    if (e & kEmit) {
      if (!BUILD || (e & kExactBit)) emit((int)(e & kLow30));
      continue;
    }
    if (BUILD && (e & kCacheNode)) {
      keep((int)(e & kLow30));
      continue;
    }
    // The four slots of a popped line: overlapping ones pushed in slot order.
    auto visit = [&](const float4 (&lo)[4], const float4 (&hi)[4]) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (!overlap(aw, bw, xyz(lo[k]), xyz(hi[k]))) continue;
        const int face = __float_as_int(hi[k].w), node = __float_as_int(lo[k].w);
        if (BUILD && face >= 0) {                 // node word under the face word: popped face first
          const bool exact = overlap(a, b, xyz(lo[k]), xyz(hi[k]));
          if (own.has(face)) {                    // not stored: only counted, when in the exact list
            if (!exact) continue;
            if (sp < S) stk[(sp++) * kTravBS] = (uint32_t)face | kEmit | kExactBit;
            else flags |= kStatusStackOverflow;
            continue;
          }
          if (sp + 1 < S) {
            stk[(sp++) * kTravBS] = (uint32_t)node | kCacheNode;
            stk[(sp++) * kTravBS] = (uint32_t)face | kEmit | (exact ? kExactBit : 0u);
          } else {
            flags |= kStatusStackOverflow;
          }
          continue;
        }
        if (!BUILD && co.stale && face >= 0 && !overlap(a, b, xyz(nlo[node]), xyz(nhi[node]))) continue;
        const uint32_t w = face >= 0 ? ((uint32_t)face | kEmit) : (uint32_t)node;
        if (sp < S) stk[(sp++) * kTravBS] = w;
        else flags |= kStatusStackOverflow;
      }
    };
    float4 lo[4], hi[4];
    // Lanes of a wavefront hold nearby queries and pop the same line about a
    // third of the time (the top of the tree): the line then comes through the
    // scalar cache, sparing the vector memory pipeline its eight loads per lane
    // (it is then moved to vector registers: testing the slots from scalar
    // registers, with the slot code duplicated for this branch, was slower).
    const uint32_t e0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)e);
    if (SLX_TRAV_SCALAR && __ballot(e != e0) == 0ull) {
      // the constant address space makes a wave-uniform load a scalar one
      typedef float f4 __attribute__((ext_vector_type(4)));
      const __attribute__((address_space(4))) f4* qs =
          (const __attribute__((address_space(4))) f4*)(quad + 8 * (size_t)e0);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const f4 l = qs[2 * k], h = qs[2 * k + 1];
        lo[k] = make_float4(l.x, l.y, l.z, l.w);
        hi[k] = make_float4(h.x, h.y, h.z, h.w);
      }
    } else {
      const float4* qn = quad + 8 * (size_t)e;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        lo[k] = qn[2 * k];
        hi[k] = qn[2 * k + 1];
      }
    }
    visit(lo, hi);
    if (co.redo && (flags & kStatusStackOverflow)) break;   // walked again by the second tier
  }
  if (BUILD) {
    pack.finish(ccnt, co.cap, nq, t, co.leaves);
    co.count[t] = (flags & kStatusStackOverflow) ? INT_MAX : ccnt;
  }
  if (co.redo && (flags & kStatusStackOverflow)) {    // the exact list may be incomplete: walk it again
    co.redo[1 + atomicAdd(&co.redo[0], 1)] = q;
    return;
  }
  cand_n[q] = kept;
  if (flags && status) atomicOr(status, flags);
  }(t0);
}

}  // namespace slx

namespace slx {
// A launch whose query count is on the device (n_active) gets at most
// knobs().active_grid workgroups (default 2048) and loops over the queries: with
// few or none active it ends in microseconds instead of dispatching a workgroup
// per possible query.
static inline int active_grid(int nq, int bs, const int32_t* n_active) {
  const int g = blocks_for(nq, bs);
  const int cap = knobs().active_grid;
  return n_active && g > cap ? cap : g;
}

int traverse_launch(const char* name, int mode, int nq, const float* q4, float padding, int n_leaves,
                    const int32_t* num, const int32_t* sorted_faces, const float* nlo4, const float* nhi4,
                    const float* quad4, const float* fmin4, const float* fmax4, const float* box6, int cap,
                    int32_t* cand, int32_t* cand_n, int32_t* status, hipStream_t st, const int32_t* roots,
                    const int32_t* q_inst, const int32_t* order, const int32_t* n_active,
                    const int32_t* qedges, const int32_t* own) {
  SLX_REQUIRE(!roots || q_inst, "slx_traverse: per-instance roots need query instance ids");
  SLX_REQUIRE(!own || !qedges, "slx_traverse: own faces belong to vertex queries");
  SLX_REQUIRE(!qedges || !roots, "slx_traverse: edge queries of batched trees are not supported");
  SLX_REQUIRE(!n_active || (order && mode == 1), "slx_traverse: a device query count needs a query list");
  SLX_REQUIRE(mode == 0 || mode == 1, "slx_traverse: mode must be 0 (cell) or 1 (triangle)");
  SLX_REQUIRE(!knobs().stale_disp || (mode == 1 && quad4 && SLX_TRAV_QUAD && !roots &&
                                      (long long)n_leaves * 8 > (long long)nq),
              "slx_traverse: boxes older than the leaves need the per-lane 4-wide walk");
  SLX_REQUIRE(nq >= 0 && cap >= 0, "slx_traverse: bad sizes");
  if (nq == 0) return SLX_OK;
  SLX_REQUIRE(q4 && cand_n && (cap == 0 || cand), "slx_traverse: null pointer");
  SLX_REQUIRE(n_leaves <= 0 || (nlo4 && nhi4), "slx_traverse: null tree");
  SLX_REQUIRE(mode == 1 || (num && sorted_faces && fmin4 && fmax4 && box6), "slx_traverse: null cell tree");
  SLX_REQUIRE(mode == 0 || n_leaves < (1 << 29), "slx_traverse: triangle tree too large for 30-bit stack words");
  if (mode == 0 && !roots && cap <= kSplitCap && n_leaves == kCells)
    SLX_LAUNCH(name, k_traverse_cell_split, dim3(blocks_for(nq, kSplitQ)), dim3(256), 0, st, nq, (const float4*)q4,
               padding, qedges, num, sorted_faces, (const float4*)nlo4, (const float4*)nhi4, (const float4*)fmin4,
               (const float4*)fmax4, box6, cap, cand, cand_n, status, order);
  else if (mode == 0)
    SLX_LAUNCH(name, k_traverse<true>, dim3(blocks_for(nq, kTravBS)), dim3(kTravBS), 0, st, nq,
               (const float4*)q4, padding, qedges, n_leaves, num, sorted_faces, (const float4*)nlo4,
               (const float4*)nhi4, (const float4*)fmin4, (const float4*)fmax4, box6, cap, cand, cand_n, status, roots,
               q_inst, order);
  // One walk per lane, unless the tree is coarse next to the queries (far fewer
  // leaves than queries: the lanes of a wavefront then overlap the same few
  // leaves and the shared walk is short).  Batched trees: lanes have different
  // roots.  The candidate lists are the same either way.
  else if (roots || !SLX_TRAV_WAVE || (long long)n_leaves * 8 > (long long)nq) {
    if (quad4 && SLX_TRAV_QUAD) {
      CacheOut co{};
      co.own = (const int4*)own;
      co.stale = knobs().stale_disp;
      SLX_LAUNCH(name, (k_traverse_quad<false, kStack>), dim3(active_grid(nq, kTravBS, n_active)), dim3(kTravBS), 0, st, nq,
                 (const float4*)q4, padding, qedges, n_leaves, (const float4*)nlo4, (const float4*)nhi4,
                 (const float4*)quad4, cap, cand, cand_n, status, roots, q_inst, order, n_active, co);
    }
    else
      SLX_LAUNCH(name, k_traverse_tri, dim3(blocks_for(nq, kTravBS)), dim3(kTravBS), 0, st, nq, (const float4*)q4,
                 padding, qedges, n_leaves, (const float4*)nlo4, (const float4*)nhi4, cap, cand, cand_n, status, roots,
                 q_inst, order, n_active);
  }
  else
    SLX_LAUNCH(name, k_traverse_tri_wave<false>, dim3(active_grid(nq, 256, n_active)), dim3(256), 0, st, nq, (const float4*)q4,
               padding, qedges, n_leaves, (const float4*)nlo4, (const float4*)nhi4, cap, cand, cand_n, status, order,
               n_active, CacheOut{});
  return check_launch("slx_traverse");
}
}  // namespace slx

namespace slx {

// ===================== candidate cache (Verlet lists over a fixed topology) =====================
//
// Within a frame the triangle tree keeps its topology; only its boxes move at the
// refits.  The walks emit leaves in one fixed order (right subtree first, i.e.
// decreasing leaf position), so a query's list is "the leaves whose box overlaps
// the query box, in that order".  A walk with the query box enlarged by a margin
// m stores the query's leaves (node ids, in walk order) and its position p0.
// Later, with the tree's vertices displaced by at most D since the snapshot the
// walk used (x = the vertices the leaf boxes are built from) and the query by d
// = |p - p0|_inf, every leaf overlapping the exact box now overlapped the
// enlarged box then, when d + D <= m (a leaf box is the padded box of its
// vertices, so it moves by at most D; the query box by d).  Filtering the stored
// leaves by the exact box then gives exactly the walk's list and order.  The test
// keeps an allowance of 2^-20 of the coordinates' magnitude for rounding.  A
// query whose condition fails, whose stored list was incomplete (more than the
// cache's capacity, or a stack overflow), or whose list may overflow, is
// appended to a list and walked in full.

__global__ void __launch_bounds__(256)
k_cand_filter(int nq, const float4* __restrict__ q4, float pad, float margin, const float4* __restrict__ nlo,
              const float4* __restrict__ nhi, int ccap, const int32_t* __restrict__ cache,
              const int32_t* __restrict__ cache_n, const float4* __restrict__ q0, const uint32_t* __restrict__ disp,
              int cap, int32_t* __restrict__ cand, int32_t* __restrict__ cand_n, int32_t* __restrict__ redo,
              const int32_t* __restrict__ order) {
  const int t = xcd_block() * 256 + threadIdx.x;
  const bool active = t < nq;
  const int q = active ? (order ? order[t] : t) : 0;
  bool ok = false;
  int n = 0, nown = 0;
  float3 p = make_float3(0.f, 0.f, 0.f);
  if (active) {
    const float4 c0 = q0[t];
    const float3 p0 = xyz(c0);
    nown = __float_as_int(c0.w);
    p = xyz(q4[q]);
    const float dq = fmaxf(fmaxf(fabsf(p.x - p0.x), fabsf(p.y - p0.y)), fabsf(p.z - p0.z));
    const float D = disp ? __uint_as_float(*disp) : 0.0f;
    const float mag = fmaxf(fmaxf(fabsf(p.x), fabsf(p.y)), fabsf(p.z)) + 2.0f * (pad + margin);
    n = cache_n[t];
    ok = n <= ccap && (dq + D) + ldexpf(mag, -20) <= margin;   // false for NaN
  }
  // the stored leaves overlapping the exact box, in order
  int cnt = 0;
  if (ok) {
    const float pp = pad * 1.0f;
    const float3 a = make_float3(p.x - pp, p.y - pp, p.z - pp), b = make_float3(p.x + pp, p.y + pp, p.z + pp);
    const int4* rows = reinterpret_cast<const int4*>(cache);
    for (int j0 = 0; j0 < n; j0 += 4) {
      const int4 r = rows[(size_t)(j0 >> 2) * nq + t];
      const int Ls[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (j0 + k >= n) break;
        const float4 lo = nlo[Ls[k]], hi = nhi[Ls[k]];
        if (!overlap(a, b, xyz(lo), xyz(hi))) continue;
        if (cnt < cap) cand[(size_t)cnt * nq + q] = __float_as_int(lo.w);
        ++cnt;
      }
    }
  }
  // queries to walk again, appended with one atomic per wavefront: a stale cache,
  // or a list that may overflow (with the nown own faces the build left out of
  // the cache, the full list has at most cnt + nown entries; the walk then
  // decides, and sets the overflow bit if it does)
  const bool need = active && (!ok || cnt + nown > cap);
  const unsigned long long m = __ballot(need);
  if (m) {
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&redo[0], __popcll(m));
    base = __shfl(base, leader, 64);
    if (need) redo[1 + base + __popcll(m & ((1ull << lane) - 1ull))] = q;
  }
  if (active && !need) cand_n[q] = cnt;
}

__global__ void k_zero_one(int32_t* __restrict__ a) {
  if (threadIdx.x == 0) a[0] = 0;
}

// x_snap = x; disp = 0
__global__ void k_tree_snapshot(int nv, const float4* __restrict__ x4, float4* __restrict__ x_snap,
                                uint32_t* __restrict__ disp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) x_snap[i] = x4[i];
  if (i == 0) disp[0] = 0u;
}

// disp = max_v |x - x_snap|_inf (disp zeroed before; non-negative floats order as
// their bits)
// (each workgroup reduces 8 vertices per thread; one atomic per workgroup, skipped
// when the stored maximum is already larger)
constexpr int kDispPer = 8;
__global__ void __launch_bounds__(256)
k_tree_disp(int nv, const float4* __restrict__ x4, const float4* __restrict__ x_snap, uint32_t* __restrict__ disp) {
  __shared__ float s_max[4];
  float d = 0.f;
  const int i0 = blockIdx.x * (256 * kDispPer) + threadIdx.x;
#pragma unroll
  for (int k = 0; k < kDispPer; ++k) {
    const int i = i0 + k * 256;
    if (i < nv) {
      const float4 x = x4[i], s0 = x_snap[i];
      float e = fmaxf(fmaxf(fabsf(x.x - s0.x), fabsf(x.y - s0.y)), fabsf(x.z - s0.z));
      if (e != e) e = __int_as_float(0x7f800000);          // NaN: no cached list stays valid
      d = fmaxf(d, e);
    }
  }
  for (int o = 32; o > 0; o >>= 1) d = fmaxf(d, __shfl_xor(d, o, 64));
  if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    d = fmaxf(fmaxf(s_max[0], s_max[1]), fmaxf(s_max[2], s_max[3]));
    if (d > 0.f && __float_as_uint(d) > __atomic_load_n(disp, __ATOMIC_RELAXED)) atomicMax(disp, __float_as_uint(d));
  }
}

__global__ void k_zero_u32(uint32_t* __restrict__ a) {
  if (threadIdx.x == 0) a[0] = 0u;
}

// first-tier quad walk with the shallow stack depth in use
template <bool BUILD>
int quad_first_tier(const char* name, int nq, const float* q4, float padding, int n_leaves, const float* nlo4,
                    const float* nhi4, const float* quad4, int cap, int32_t* cand, int32_t* cand_n, int32_t* status,
                    hipStream_t st, const int32_t* order, const CacheOut& co, const int32_t* qedges,
                    const int32_t* n_active, const int32_t* roots = nullptr, const int32_t* q_inst = nullptr) {
  const dim3 g(active_grid(nq, kTravBS, n_active)), b(kTravBS);
#define SLX_QUAD_TIER(D)                                                                                         \
  SLX_LAUNCH(name, (k_traverse_quad<BUILD, D>), g, b, 0, st, nq, (const float4*)q4, padding, qedges, n_leaves,    \
             (const float4*)nlo4, (const float4*)nhi4, (const float4*)quad4, cap, cand, cand_n, status, roots,    \
             q_inst, order, n_active, co)
  const int depth = knobs().walk_stack;
  if (depth <= 16) SLX_QUAD_TIER(16);
  else if (depth <= 24) SLX_QUAD_TIER(24);
  else if (depth <= 32) SLX_QUAD_TIER(32);
  else SLX_QUAD_TIER(kStack);
#undef SLX_QUAD_TIER
  return SLX_OK;
}

int traverse_cached(const char* name, int kind, int nq, const float* q4, float padding, int n_leaves,
                    const float* nlo4, const float* nhi4, const float* quad4, const slx_cand_cache* c,
                    const uint32_t* disp, int cap, int32_t* cand, int32_t* cand_n, int32_t* status, hipStream_t st,
                    const int32_t* order, const int32_t* qedges, bool prezeroed, const int32_t* n_active,
                    const int32_t* roots, const int32_t* q_inst, const int32_t* own, const char* redo_name) {
  SLX_REQUIRE(kind >= 0 && kind <= 2, "slx_traverse_cached: kind must be 0 (walk), 1 (build) or 2 (filter)");
  SLX_REQUIRE(!own || !qedges, "slx_traverse_cached: own faces belong to vertex queries");
  SLX_REQUIRE(!qedges || kind == 0, "slx_traverse_cached: edge queries are walked, not cached");
  SLX_REQUIRE(!n_active || (kind == 0 && order), "slx_traverse_cached: a device query count needs a walk and a list");
  SLX_REQUIRE(c && c->redo && (kind == 0 || (c->cap >= 1 && c->leaves && c->count && c->q0)),
              "slx_traverse_cached: bad cache");
  SLX_REQUIRE(nq >= 0 && cap >= 0 && c->margin >= 0.f, "slx_traverse_cached: bad sizes");
  if (nq == 0) return SLX_OK;
  SLX_REQUIRE(q4 && cand_n && (cap == 0 || cand), "slx_traverse_cached: null pointer");
  SLX_REQUIRE(n_leaves < (1 << 29), "slx_traverse_cached: tree too large for 30-bit stack words");
  if (n_leaves <= 0) return traverse_launch(name, 1, nq, q4, padding, n_leaves, nullptr, nullptr, nlo4, nhi4, quad4,
                                            nullptr, nullptr, nullptr, cap, cand, cand_n, status, st, nullptr,
                                            nullptr, order, n_active, qedges);
  SLX_REQUIRE(nlo4 && nhi4, "slx_traverse_cached: null tree");
  SLX_REQUIRE(!roots || (q_inst && !qedges), "slx_traverse_cached: per-instance roots need query instance ids");
  // (trees of batched instances are walked from each query's instance root: per lane)
  const bool wave = !roots && SLX_TRAV_WAVE && (long long)n_leaves * 8 <= (long long)nq;
  SLX_REQUIRE(wave || quad4, "slx_traverse_cached: the triangle tree needs its 4-wide node array");
  SLX_REQUIRE(!knobs().stale_disp || (kind != 1 && !wave && !roots),
              "slx_traverse_cached: boxes older than the leaves need the per-lane walk, and no cache build");
  if (kind == 0 && wave)                      // the wavefront walk keeps its stack in registers
    return traverse_launch(name, 1, nq, q4, padding, n_leaves, nullptr, nullptr, nlo4, nhi4, quad4, nullptr,
                           nullptr, nullptr, cap, cand, cand_n, status, st, nullptr, nullptr, order, n_active, qedges);
  if (!prezeroed) SLX_LAUNCH("fill", k_zero_one, dim3(1), dim3(64), 0, st, c->redo);
  // dense point queries of a fine tree (about as many queries as leaves, in a
  // coherent order): the 4-wide wavefront walk when knobs().walk_wave is on (off
  // by default: on the 1M-vertex drape its cache build took 1018 us against the
  // per-lane walk's 366 us -- the shared walk is a long chain of dependent loads)
  const bool quad_wave =
      !wave && !roots && knobs().walk_wave && !knobs().stale_disp && !qedges && !n_active && (long long)nq * 4 >= n_leaves;
  int rc = SLX_OK;
  if (kind == 1) {
    const CacheOut co{padding + c->margin, c->cap, c->leaves, c->count, (float4*)c->q0, c->redo, (const int4*)own};
    if (wave)
      SLX_LAUNCH("cache_build_wave", k_traverse_tri_wave<true>, dim3(blocks_for(nq, 256)), dim3(256), 0, st, nq,
                 (const float4*)q4, padding, nullptr, n_leaves, (const float4*)nlo4, (const float4*)nhi4, cap, cand,
                 cand_n, status, order, nullptr, co);
    else if (quad_wave)
      SLX_LAUNCH("cache_build", k_traverse_quad_wave<true>, dim3(blocks_for(nq, 256)), dim3(256), 0, st, nq,
                 (const float4*)q4, padding, nullptr, n_leaves, (const float4*)nlo4, (const float4*)nhi4,
                 (const float4*)quad4, cap, cand, cand_n, status, order, nullptr, co);
    else
      rc = quad_first_tier<true>("cache_build", nq, q4, padding, n_leaves, nlo4, nhi4, quad4, cap, cand, cand_n,
                                 status, st, order, co, nullptr, nullptr, roots, q_inst);
  } else if (kind == 2) {
    SLX_LAUNCH(name, k_cand_filter, dim3(blocks_for(nq, 256)), dim3(256), 0, st, nq, (const float4*)q4, padding,
               c->margin, (const float4*)nlo4, (const float4*)nhi4, c->cap, c->leaves, c->count,
               (const float4*)c->q0, disp, cap, cand, cand_n, c->redo, order);
  } else {
    CacheOut co{};
    co.redo = c->redo;
    co.own = (const int4*)own;
    co.stale = knobs().stale_disp;
    if (quad_wave)
      SLX_LAUNCH(name, k_traverse_quad_wave<false>, dim3(blocks_for(nq, 256)), dim3(256), 0, st, nq,
                 (const float4*)q4, padding, nullptr, n_leaves, (const float4*)nlo4, (const float4*)nhi4,
                 (const float4*)quad4, cap, cand, cand_n, status, order, nullptr, co);
    else
      rc = quad_first_tier<false>(name, nq, q4, padding, n_leaves, nlo4, nhi4, quad4, cap, cand, cand_n, status,
                                  st, order, co, qedges, n_active, roots, q_inst);
  }
  if (rc) return rc;
  // the queries the first tier could not serve: a full walk over the list built on the device
  return traverse_launch(redo_name ? redo_name : "walk_redo", 1, nq, q4, padding, n_leaves, nullptr, nullptr, nlo4,
                         nhi4, quad4, nullptr,
                         nullptr, nullptr, cap, cand, cand_n, status, st, roots, q_inst, c->redo + 1, c->redo,
                         qedges, own);
}

}  // namespace slx

extern "C" int slx_traverse_cached(int kind, int nq, const float* q4, float padding, int n_leaves, const float* nlo4,
                                   const float* nhi4, const float* quad4, const slx_cand_cache* cache,
                                   const uint32_t* disp, int cap, int32_t* cand, int32_t* cand_n, int32_t* status,
                                   const int32_t* order, slx_stream_t stream) {
  return traverse_cached("traverse_cached", kind, nq, q4, padding, n_leaves, nlo4, nhi4, quad4, cache, disp, cap,
                         cand, cand_n, status, S(stream), order, nullptr);
}

extern "C" int slx_set_active_grid(int blocks) {
  SLX_REQUIRE(blocks >= 1 && blocks <= 65535, "slx_set_active_grid: 1..65535 workgroups");
  knob_defaults().active_grid = blocks;
  return SLX_OK;
}

extern "C" int slx_set_walk_stack(int depth) {
  SLX_REQUIRE(depth == 16 || depth == 24 || depth == 32 || depth == 64, "slx_set_walk_stack: 16, 24, 32 or 64");
  knob_defaults().walk_stack = depth;
  return SLX_OK;
}

extern "C" int slx_tree_snapshot(int nv, const float* x4, float* x_snap, uint32_t* disp, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0 && disp && (nv == 0 || (x4 && x_snap)), "slx_tree_snapshot: bad arguments");
  SLX_LAUNCH("tree_snapshot", k_tree_snapshot, dim3(blocks_for(nv > 0 ? nv : 1, 256)), dim3(256), 0, S(stream), nv,
             (const float4*)x4, (float4*)x_snap, disp);
  return check_launch("slx_tree_snapshot");
}

namespace slx {
int tree_displacement(int nv, const float* x4, const float* x_snap, uint32_t* disp, bool zero, hipStream_t st) {
  SLX_REQUIRE(nv >= 0 && disp && (nv == 0 || (x4 && x_snap)), "slx_tree_displacement: bad arguments");
  if (zero) SLX_LAUNCH("fill", k_zero_u32, dim3(1), dim3(64), 0, st, disp);
  if (nv > 0)
    SLX_LAUNCH("tree_disp", k_tree_disp, dim3(blocks_for(nv, 256 * kDispPer)), dim3(256), 0, st, nv,
               (const float4*)x4, (const float4*)x_snap, disp);
  return check_launch("slx_tree_displacement");
}
}  // namespace slx

extern "C" int slx_tree_displacement(int nv, const float* x4, const float* x_snap, uint32_t* disp,
                                     slx_stream_t stream) {
  return tree_displacement(nv, x4, x_snap, disp, true, S(stream));
}

extern "C" int slx_traverse(int mode, int nq, const float* q4, float padding, int n_leaves, const int32_t* num,
                            const int32_t* sorted_faces, const float* nlo4, const float* nhi4, const float* quad4,
                            const float* fmin4, const float* fmax4, const float* box6, int cap, int32_t* cand,
                            int32_t* cand_n, int32_t* status, slx_stream_t stream) {
  return traverse_launch(mode == 0 ? "traverse_cell" : "traverse_tri", mode, nq, q4, padding, n_leaves, num,
                         sorted_faces, nlo4, nhi4, quad4, fmin4, fmax4, box6, cap, cand, cand_n, status, S(stream),
                         nullptr, nullptr, nullptr);
}

extern "C" int slx_traverse_edges(int mode, int ne, const int32_t* edges, const float* x4, float padding,
                                  int n_leaves, const int32_t* num, const int32_t* sorted_faces, const float* nlo4,
                                  const float* nhi4, const float* quad4, const float* fmin4, const float* fmax4,
                                  const float* box6, int cap, int32_t* cand, int32_t* cand_n, int32_t* status,
                                  const int32_t* order, int32_t* redo, slx_stream_t stream) {
  SLX_REQUIRE(ne == 0 || (edges && x4), "slx_traverse_edges: null edges or positions");
  if (mode == 1 && redo && n_leaves > 0) {
    slx_cand_cache walk{};
    walk.redo = redo;
    return traverse_cached("traverse_ee", 0, ne, x4, padding, n_leaves, nlo4, nhi4, quad4, &walk, nullptr, cap, cand,
                           cand_n, status, S(stream), order, edges, false);
  }
  return traverse_launch(mode == 0 ? "traverse_ee_cell" : "traverse_ee", mode, ne, x4, padding, n_leaves, num,
                         sorted_faces, nlo4, nhi4, quad4, fmin4, fmax4, box6, cap, cand, cand_n, status, S(stream),
                         nullptr, nullptr, order, nullptr, edges);
}

// ===================== batched instances: one tree, one subtree per instance =====================
//
// Instances are packed CSR-style: vertices of instance i are [vert_base[i],
// vert_base[i+1]), faces [face_base[i], face_base[i+1]) (framework/meshio/
// concat.py:11-58 packs meshes the same way).  Each face gets a Morton code in its
// own instance's box; a stable sort by code followed by a stable sort by instance
// orders faces by (instance, code, index), the composite 64-bit key
// inst<<54 | code<<24 | index.  The Karras tree over those keys contains, for
// every instance, exactly one node whose range is that instance's faces (the
// keys sharing the instance prefix); roots[i] is that node, and a query of
// instance i starts its traversal there, so instances never see each other.
namespace slx {

constexpr int kSegInstBits = 10;      // up to 1024 instances
constexpr int kSegIndexBits = 24;     // up to 16M faces

__global__ void __launch_bounds__(256)
k_seg_box(const int32_t* __restrict__ vert_base, const float4* __restrict__ x4, float pad, float* __restrict__ box) {
  __shared__ float s[6][256];
  const int inst = blockIdx.x;
  const int b = vert_base[inst], e = vert_base[inst + 1];
  float lo[3] = {1e5f, 1e5f, 1e5f}, hi[3] = {-1e5f, -1e5f, -1e5f};
  for (int i = b + (int)threadIdx.x; i < e; i += 256) {
    float4 p = x4[i];
    lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
    hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
  }
  for (int c = 0; c < 3; ++c) { s[c][threadIdx.x] = lo[c]; s[3 + c][threadIdx.x] = hi[c]; }
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = fminf(s[c][threadIdx.x], s[c][threadIdx.x + o]);
        s[3 + c][threadIdx.x] = fmaxf(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + o]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 3) box[6 * inst + threadIdx.x] = s[threadIdx.x][0] - pad * 1.0f;
  else if (threadIdx.x < 6) box[6 * inst + threadIdx.x] = s[threadIdx.x][0] + pad * 1.0f;
}

__global__ void k_face_morton_seg(int nf, const float4* __restrict__ fmin, const float4* __restrict__ fmax,
                                  const int32_t* __restrict__ face_inst, const float* __restrict__ box,
                                  uint32_t* __restrict__ codes, uint32_t* __restrict__ codes_keep,
                                  int32_t* __restrict__ ids) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const float* bx = box + 6 * face_inst[f];
  float4 a = fmin[f], b = fmax[f];
  float3 c = make_float3(0.5f * (a.x + b.x), 0.5f * (a.y + b.y), 0.5f * (a.z + b.z));
  uint32_t code = morton_in_cube(c, make_float3(bx[0], bx[1], bx[2]), make_float3(bx[3], bx[4], bx[5]));
  codes[f] = code;
  codes_keep[f] = code;
  ids[f] = f;
}

__global__ void k_gather_inst(int nf, const int32_t* __restrict__ ids, const int32_t* __restrict__ face_inst,
                              uint32_t* __restrict__ key) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nf) key[k] = (uint32_t)face_inst[ids[k]];
}

__global__ void k_gather_codes(int nf, const int32_t* __restrict__ ids, const uint32_t* __restrict__ codes_keep,
                               const int32_t* __restrict__ face_inst, uint32_t* __restrict__ codes,
                               int32_t* __restrict__ sinst) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nf) return;
  int f = ids[k];
  codes[k] = codes_keep[f];
  sinst[k] = face_inst[f];
}

// composite key; the tie-break index is instance-local so every instance's
// subtree has exactly the shape of a tree built over that instance alone
__device__ __forceinline__ uint64_t seg_key(const uint32_t* codes, const int32_t* sinst, const int32_t* fbase, int i) {
  const int inst = sinst[i];
  return ((uint64_t)(uint32_t)inst << (64 - kSegInstBits)) | ((uint64_t)codes[i] << kSegIndexBits) |
         (uint32_t)(i - fbase[inst]);
}
__device__ __forceinline__ int seg_prefix(const uint32_t* codes, const int32_t* sinst, const int32_t* fbase, int n,
                                          int i, int j) {
  if (j < 0 || j >= n) return -1;
  return __clzll(seg_key(codes, sinst, fbase, i) ^ seg_key(codes, sinst, fbase, j));
}

// Karras 2012 over the composite keys (distinct by construction)
__global__ void k_karras_seg(int n, const uint32_t* __restrict__ codes, const int32_t* __restrict__ sinst,
                             const int32_t* __restrict__ fbase, int32_t* __restrict__ ca, int32_t* __restrict__ cb, int32_t* __restrict__ parent,
                             int32_t* __restrict__ rl, int32_t* __restrict__ rr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) parent[0] = -1;
  if (i >= n - 1) return;
  int dl = seg_prefix(codes, sinst, fbase, n, i, i - 1), dr = seg_prefix(codes, sinst, fbase, n, i, i + 1);
  int d = (dr < dl) ? -1 : 1;
  int dmin = (dr < dl) ? dr : dl;
  int hi = 2;
  while (seg_prefix(codes, sinst, fbase, n, i, i + hi * d) > dmin) hi *= 2;
  int len = 0;
  for (int t = hi / 2; t > 0; t /= 2)
    if (seg_prefix(codes, sinst, fbase, n, i, i + (len + t) * d) > dmin) len += t;
  int j = i + len * d;
  int first = d > 0 ? i : j, last = d > 0 ? j : i;
  int node_prefix = seg_prefix(codes, sinst, fbase, n, first, last);
  int split = first, step = last - first;
  do {
    step = (step + 1) / 2;
    int c = split + step;
    if (c < last && seg_prefix(codes, sinst, fbase, n, first, c) > node_prefix) split = c;
  } while (step > 1);
  int a = (split == first) ? split + (n - 1) : split;
  int b = (split + 1 == last) ? split + 1 + (n - 1) : split + 1;
  ca[i] = a;
  cb[i] = b;
  parent[a] = i;
  parent[b] = i;
  rl[i] = first;
  rr[i] = last;
}

// roots[i]: the node covering exactly instance i's sorted faces
__global__ void k_seg_roots_leaf(int n_inst, int n, const int32_t* __restrict__ face_base, int32_t* __restrict__ roots) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst) return;
  int c = face_base[i + 1] - face_base[i];
  roots[i] = c == 1 ? face_base[i] + (n - 1) : -1;
}
__global__ void k_seg_roots_internal(int n, const int32_t* __restrict__ rl, const int32_t* __restrict__ rr,
                                     const int32_t* __restrict__ sinst, const int32_t* __restrict__ face_base,
                                     int32_t* __restrict__ roots) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int l = rl[i], r = rr[i], inst = sinst[l];
  if (l == face_base[inst] && r == face_base[inst + 1] - 1) roots[inst] = i;
}

}  // namespace slx

extern "C" size_t slx_seg_build_ws_bytes(int nf, int n_inst) {
  size_t n = (size_t)(nf > 0 ? nf : 1);
  (void)n_inst;
  return 16 * n + radix_sort_ws_bytes(nf) + 2048;
}

extern "C" int slx_seg_build(int nv, int nf, int n_inst, const int32_t* faces, const float* x4, float padding,
                             const int32_t* vert_base, const int32_t* face_base, const int32_t* face_inst,
                             float* fmin4, float* fmax4, float* box6, int32_t* codes, int32_t* sorted_faces,
                             int32_t* sorted_inst, int32_t* child_a, int32_t* child_b, int32_t* parent,
                             int32_t* range_l, int32_t* range_r, int32_t* roots, int32_t* sched, float* nlo4,
                             float* nhi4, float* quad4, void* ws, size_t ws_bytes, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0 && nf >= 0 && n_inst >= 1 && n_inst <= (1 << kSegInstBits), "slx_seg_build: bad sizes");
  SLX_REQUIRE(nf < (1 << kSegIndexBits), "slx_seg_build: too many faces for 24-bit indices");
  if (nf == 0) return SLX_OK;
  SLX_REQUIRE(faces && x4 && vert_base && face_base && face_inst && fmin4 && fmax4 && box6 && codes && sorted_faces &&
                  sorted_inst && child_a && child_b && parent && range_l && range_r && roots && sched && nlo4 &&
                  nhi4 && ws,
              "slx_seg_build: null pointer");
  if (ws_bytes < slx_seg_build_ws_bytes(nf, n_inst)) return fail(SLX_EWS, "slx_seg_build: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  char* p = (char*)ws;
  uint32_t* keep = (uint32_t*)p; p += 4 * (size_t)nf;
  uint32_t* key2 = (uint32_t*)p; p += 4 * (size_t)nf;
  p = (char*)ws + 16 * (size_t)nf + 1024;
  void* sort_ws = p;
  size_t sort_bytes = ws_bytes - (16 * (size_t)nf + 1024);
  int rc = slx_face_aabbs(nf, faces, x4, padding, fmin4, fmax4, stream);
  if (rc) return rc;
  SLX_LAUNCH("seg_box", k_seg_box, dim3(n_inst), dim3(256), 0, st, vert_base, (const float4*)x4, padding, box6);
  SLX_LAUNCH("face_morton", k_face_morton_seg, dim3(blocks_for(nf, 256)), dim3(256), 0, st, nf, (const float4*)fmin4,
             (const float4*)fmax4, face_inst, box6, (uint32_t*)codes, keep, sorted_faces);
  rc = radix_sort_pairs(nf, nullptr, (uint32_t*)codes, sorted_faces, 30, sort_ws, sort_bytes, st);
  if (rc) return rc;
  SLX_LAUNCH("gather", k_gather_inst, dim3(blocks_for(nf, 256)), dim3(256), 0, st, nf, sorted_faces, face_inst, key2);
  rc = radix_sort_pairs(nf, nullptr, key2, sorted_faces, kSegInstBits, sort_ws, sort_bytes, st);
  if (rc) return rc;
  SLX_LAUNCH("gather", k_gather_codes, dim3(blocks_for(nf, 256)), dim3(256), 0, st, nf, sorted_faces, keep, face_inst,
             (uint32_t*)codes, sorted_inst);
  SLX_LAUNCH("karras", k_karras_seg, dim3(blocks_for(nf, 256)), dim3(256), 0, st, nf, (const uint32_t*)codes,
             sorted_inst, face_base, child_a, child_b, parent, range_l, range_r);
  SLX_LAUNCH("seg_roots", k_seg_roots_leaf, dim3(blocks_for(n_inst, 256)), dim3(256), 0, st, n_inst, nf, face_base,
             roots);
  if (nf > 1)
    SLX_LAUNCH("seg_roots", k_seg_roots_internal, dim3(blocks_for(nf - 1, 256)), dim3(256), 0, st, nf, range_l,
               range_r, sorted_inst, face_base, roots);
  rc = check_launch("slx_seg_build");
  if (rc) return rc;
  rc = refit_schedule(nf, child_b, parent, range_l, range_r, sched, st, false);
  if (rc) return rc;
  LeafSrc src{};
  src.fmin = (const float4*)fmin4;
  src.fmax = (const float4*)fmax4;
  src.sorted = sorted_faces;
  src.quad = (float4*)quad4;
  return refit<kLeafFaceBoxes>("tri_refit", nf, sched, src, child_a, child_b, parent, nullptr, (float4*)nlo4,
                               (float4*)nhi4, st);
}

extern "C" int slx_traverse_seg(int nq, const float* q4, const int32_t* q_inst, const int32_t* roots, float padding,
                                int nf, const float* nlo4, const float* nhi4, const float* quad4, int cap,
                                int32_t* cand, int32_t* cand_n, int32_t* status, slx_stream_t stream) {
  SLX_REQUIRE(roots && q_inst, "slx_traverse_seg: roots and q_inst required");
  return traverse_launch("traverse_seg", 1, nq, q4, padding, nf, nullptr, nullptr, nlo4, nhi4, quad4, nullptr,
                         nullptr, nullptr, cap, cand, cand_n, status, S(stream), roots, q_inst, nullptr);
}
