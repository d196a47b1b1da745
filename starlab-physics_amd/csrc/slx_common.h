// Shared device helpers for the gfx950 XPBD cloth kernels.
//
// Floating-point op order in every formula is written out explicitly and the
// library is compiled with -ffp-contract=off, so each kernel performs the same
// IEEE single-precision operations, in the same order, as the serial
// restatement of the reference step; division and sqrt are the correctly
// rounded HIP defaults.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/slx.h"

namespace slx {

constexpr int kWave = 64;
constexpr int kGrid = 32;                       // cell LBVH resolution per axis
constexpr int kCells = kGrid * kGrid * kGrid;   // 32768

// status word bits (device int32 written with atomicOr)
constexpr int kStatusCandOverflow = SLX_STATUS_CAND_OVERFLOW;
constexpr int kStatusStackOverflow = SLX_STATUS_STACK_OVERFLOW;
constexpr int kStatusCellRange = SLX_STATUS_CELL_RANGE;
constexpr int kStatusRecOverflow = SLX_STATUS_REC_OVERFLOW;
constexpr int kStatusPairOverflow = SLX_STATUS_PAIR_OVERFLOW;

// The dispatcher hands workgroups to the 8 XCDs round-robin by workgroup id,
// and each XCD has its own L2.  xcd_block() renumbers a launch's workgroups so
// that each XCD runs one contiguous range of logical blocks: blocks adjacent in
// the processing order (queries close in space, sharing tree lines, faces and
// vertices; neighbouring cloth rows) then share one L2 instead of eight.  It is
// a bijection of [0, gridDim.x) (XCD x takes gridDim/8 blocks, one more for the
// first gridDim % 8 XCDs).  Kernels that use it compute nothing from the block
// order, so their results do not depend on it.  (Not for a launch whose active
// work is a prefix of its grid, e.g. a count known only on the device: grouped
// by XCD, that prefix would run on the first XCDs alone.)
#ifndef SLX_XCD_MAP
#define SLX_XCD_MAP 1
#endif
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
#if SLX_XCD_MAP
  const int g = (int)gridDim.x, b = (int)blockIdx.x;
  const int x = b % kXcds, tall = g / kXcds, extra = g % kXcds;
  return x * tall + min(x, extra) + b / kXcds;
#else
  return (int)blockIdx.x;
#endif
}

__device__ __forceinline__ float dot3(float3 a, float3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 mul3(float s, float3 a) { return make_float3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 xyz(float4 a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float3 ld3(const float* p, int i) {
  return make_float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

// float -> int32 truncation: NaN -> INT32_MIN, saturating (host oracle does the same)
__device__ __forceinline__ int32_t trunc_i32(float v) {
  if (!(v > -2147483648.0f)) return INT32_MIN;
  if (v >= 2147483648.0f) return INT32_MAX;
  return (int32_t)v;
}

__device__ __forceinline__ bool overlap(float3 amin, float3 amax, float3 bmin, float3 bmax) {
  return amin.x <= bmax.x && amax.x >= bmin.x && amin.y <= bmax.y && amax.y >= bmin.y &&
         amin.z <= bmax.z && amax.z >= bmin.z;
}

inline int blocks_for(long long n, int bs) { return (int)((n + bs - 1) / bs); }

// computeAABB_faces, meshTaichiWrapper.py:350-361: componentwise min/max of the
// three corners, minus/plus the padding
__device__ __forceinline__ void corner_box(float3 a, float3 b, float3 c, float pad, float3& lo, float3& hi) {
  const float p = pad * 1.0f;
  lo = make_float3(fminf(fminf(a.x, b.x), c.x) - p, fminf(fminf(a.y, b.y), c.y) - p, fminf(fminf(a.z, b.z), c.z) - p);
  hi = make_float3(fmaxf(fmaxf(a.x, b.x), c.x) + p, fmaxf(fmaxf(a.y, b.y), c.y) + p, fmaxf(fmaxf(a.z, b.z), c.z) + p);
}
__device__ __forceinline__ void face_box(const int32_t* __restrict__ faces, const float4* __restrict__ x4, int f,
                                         float pad, float3& lo, float3& hi) {
  corner_box(xyz(x4[faces[3 * f]]), xyz(x4[faces[3 * f + 1]]), xyz(x4[faces[3 * f + 2]]), pad, lo, hi);
}

// Per-vertex contact-delta buckets (position and velocity passes share them):
// bcnt[v] claimed slots, brec[v * bucket + i] = (delta, order key bits),
// self_acc[v] = (deltas the vertex's own first-set query summed in order, count).
struct ContactWs {
  int32_t* bcnt;
  float4* self_acc;
  float4* brec;
};

inline size_t contact_ws_layout(int nv, int bucket, char* base, ContactWs* w) {
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t at = o;
    o += (bytes + 255) & ~(size_t)255;
    return base ? base + at : nullptr;
  };
  ContactWs t;
  t.bcnt = (int32_t*)take(4 * (size_t)nv);
  t.self_acc = (float4*)take(16 * (size_t)nv);
  t.brec = (float4*)take(16 * (size_t)nv * (size_t)(bucket > 0 ? bucket : 1));
  if (w) *w = t;
  return o;
}

// acc + the n bucket records r[0..n) added in increasing order key (.w bits);
// keys are distinct, so the result is that of a serial scatter in key order
__device__ __forceinline__ float3 bucket_sum(float3 acc, const float4* __restrict__ r, int n) {
  uint32_t last = 0;
  for (int i = 0; i < n; ++i) {
    uint32_t best = 0xffffffffu;
    int bj = 0;
    for (int j = 0; j < n; ++j) {
      const uint32_t o = __float_as_uint(r[j].w);
      if ((i == 0 || o > last) && o < best) {
        best = o;
        bj = j;
      }
    }
    last = best;
    const float4 d = r[bj];
    acc = make_float3(acc.x + d.x, acc.y + d.y, acc.z + d.z);
  }
  return acc;
}

// error reporting (thread-local message; see abi.hip)
int fail(int code, const char* fmt, ...);
int check_launch(const char* what);

// Jacobi spring pass launcher (xpbd.hip): pred = compute_y fused in (reads x4, v4),
// final = velocity/position update fused in (writes x_out, v_out)
int spring_launch(bool pred, bool final, int nv, const float4* x4, const float4* v4, const float* fixed,
                  const float4* y_in, float4* y_out, float4* x_out, float4* v_out, const int32_t* slice_off,
                  const int32_t* slice_width, const int32_t* entries, float k_s, float k_b, float dt, float gx,
                  float gy, float gz, float damping, hipStream_t st);

// one Gauss-Seidel iteration in place on y4 (gs.hip); color_off_host is a host array
int gs_pass(int nv, const float4* x4, const float* fixed, float4* y4, const int32_t* slice_off,
            const int32_t* slice_width, const int32_t* entries, const int4* gs_edges, const int32_t* color_off_host,
            int n_colors, float k_s, float k_b, hipStream_t st);

// stack traversal launcher (lbvh.hip); name is the kernel timer label
int traverse_launch(const char* name, int mode, int nq, const float* q4, float padding, int n_leaves,
                    const int32_t* num, const int32_t* sorted_faces, const float* nlo4, const float* nhi4,
                    const float* quad4, const float* fmin4, const float* fmax4, const float* box6, int cap,
                    int32_t* cand, int32_t* cand_n, int32_t* status, hipStream_t st, const int32_t* roots = nullptr,
                    const int32_t* q_inst = nullptr, const int32_t* order = nullptr,
                    const int32_t* n_active = nullptr, const int32_t* qedges = nullptr,
                    const int32_t* own = nullptr);

// own (nullable, [nq][8] face ids, -1 padded): each vertex query's own faces,
// counted in its list's length but not stored (the 4-wide walks and the cache
// filter; the other walks store them: either list gives the same contacts)

// two-tier traversal of a triangle tree with a candidate cache (lbvh.hip): kind 0
// = walk, 1 = walk with the enlarged box filling the cache, 2 = filter the cache;
// queries the first tier cannot serve are walked again with the deep stack
// (timer label redo_name, default "walk_redo")
int traverse_cached(const char* name, int kind, int nq, const float* q4, float padding, int n_leaves,
                    const float* nlo4, const float* nhi4, const float* quad4, const slx_cand_cache* cache,
                    const uint32_t* disp, int cap, int32_t* cand, int32_t* cand_n, int32_t* status, hipStream_t st,
                    const int32_t* order, const int32_t* qedges = nullptr, bool prezeroed = false,
                    const int32_t* n_active = nullptr, const int32_t* roots = nullptr,
                    const int32_t* q_inst = nullptr, const int32_t* own = nullptr,
                    const char* redo_name = nullptr);

// Counters a kernel zeroes for the next pass (thread i < n of block 0 writes *p[i] = 0).
struct ZeroList {
  int32_t* p[6];
  int n;
};

// The collision and velocity passes as slx_frame runs them (narrow.hip,
// velocity.hip): bcnt_zeroed -- the buckets' counters are known to be zero (every
// apply kernel leaves them so); `after` -- counters the apply kernel zeroes for the
// next pass (the walks' redo lists); set0 / before_apply: the first query set
// launched on its own (1) or left to such a launch (2, joined by before_apply).
// compute_velocity fused into the apply of a substep's last pass (its arithmetic,
// on the final y): v = (1 - damping) * fixed * (y - x) / dt, v.w = fixed
struct VelFuse {
  const float4* x4;
  float4* v4;
  float damping, dt;
};
int collision_pass_impl(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st, float* y4,
                        const float* xs4, const float* fixed, const float* m_inv, float dhat, float k_col, int cap,
                        const int32_t* c_vt_st, const int32_t* n_vt_st, const int32_t* c_vt_dy,
                        const int32_t* n_vt_dy, const int32_t* c_tv_st, const int32_t* n_tv_st, int32_t* pair_num,
                        int pair_cap, int32_t* pair_cnt, float* pair_rec, const slx_edge_sets* ee, int rec_cap,
                        int32_t* status, void* ws, size_t ws_bytes, hipStream_t st, bool bcnt_zeroed,
                        const ZeroList& after, int set0 = 0, hipEvent_t before_apply = nullptr,
                        const VelFuse* vel = nullptr);
int velocity_pass_impl(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st, float* v4,
                       const float* vs4, const float* fixed, const float* m_inv, const float* m_inv_st, float mu,
                       int pair_cap, const int32_t* pair_cnt, const float* pair_rec, const slx_edge_sets* ee,
                       float* x4, float dt, int rec_cap, int32_t* status, void* ws, size_t ws_bytes, hipStream_t st,
                       bool bcnt_zeroed);
// zero the buckets' counters (workspace ws) and the listed counters in one launch
int frame_zero(int nv, int rec_cap, void* ws, const ZeroList& zl, hipStream_t st);
// slx_tri_refit, recording after_sub (nullable) once the subtree kernel is queued.
// lines_only: the tree is read only through its 4-wide lines, its leaves, its top
// nodes and its subtree roots with their children (single trees whose walks all
// use the lines): the other internal nodes' binary boxes are left as they were.
int tri_refit_impl(int nf, const int32_t* faces, const float* x4, float padding, const int32_t* sorted_faces,
                   const int32_t* child_a, const int32_t* child_b, const int32_t* parent, const int32_t* sched,
                   float* nlo4, float* nhi4, float* quad4, hipStream_t st, hipEvent_t after_sub,
                   bool lines_only = false);
// The leaves' boxes of a triangle tree from x (the refit's first phase alone): the
// internal boxes keep the positions of the last refit.
int tri_leaf_boxes_impl(int nf, const int32_t* faces, const float* x4, float padding, const int32_t* sorted_faces,
                        float* nlo4, float* nhi4, hipStream_t st);
// slx_tree_displacement; zero = false keeps disp as it is before the max (a
// snapshot taken in the same stream order has zeroed it)
int tree_displacement(int nv, const float* x4, const float* x_snap, uint32_t* disp, bool zero, hipStream_t st);

// edge-edge phase (edges.hip): argument checks; the vertex-triangle bucket
// records folded into self_acc, then the edge-edge deltas into the emptied buckets
int ee_check(const slx_edge_sets* ee, bool velocity);
int ee_position_launch(int nv, const slx_edge_sets* ee, const float4* y4, const float4* xs4, const float* fixed,
                       const float* m_inv, float dhat, float k, const ContactWs& w, int bucket, int32_t* status,
                       hipStream_t st);
int ee_velocity_launch(int nv, const slx_edge_sets* ee, const float4* v4, const float4* vs4, const float* fixed,
                       const float* m_inv, const ContactWs& w, int bucket, int32_t* status, hipStream_t st);

// the edge gate (edges.hip): the static maxima once per frame, then per pass the
// dynamic maxima and the lists of edges whose pairs could act (null list: the set
// is not gated)
int ee_gate_static(const slx_edge_sets* ee, int nes, const float4* xs4, hipStream_t st);
int ee_gate(const slx_edge_sets* ee, int nv, const float4* x4, const float4* y4, float pad, bool new_x,
            int32_t* live_st, int32_t* live_dy, hipStream_t st);

// Launch knobs (results never depend on them): the library defaults, set by
// the slx_set_* entry points, and the values an slx_frame call takes from its
// descriptor for that call on the calling thread (KnobScope).
struct Knobs {
  int spring_block, spring_vpt, walk_stack, active_grid, refit_sub, refit_top, walk_wave;
  // (frame-internal) the dynamic tree's displacement bound when its internal boxes
  // are older than its leaf boxes: per-lane walks then enlarge their box by it and
  // test leaves by their current boxes (null: boxes current)
  const uint32_t* stale_disp;
};
Knobs& knob_defaults();
const Knobs& knobs();
class KnobScope {
 public:
  explicit KnobScope(const Knobs& k);
  ~KnobScope();

 private:
  Knobs k_;
  const Knobs* prev_;
};

// optional per-kernel event timer (slx_timer_enable); no-ops when disabled
void timer_begin(const char* name, hipStream_t s, int* slot);
void timer_end(int slot, hipStream_t s);

}  // namespace slx

// Launch with optional event timing around the kernel, on the same stream.
#define SLX_LAUNCH(name, kernel, grid, block, shmem, stream, ...)            \
  do {                                                                       \
    int _slx_ts;                                                             \
    ::slx::timer_begin(name, stream, &_slx_ts);                              \
    hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);     \
    ::slx::timer_end(_slx_ts, stream);                                       \
  } while (0)

#define SLX_REQUIRE(cond, ...)                          \
  do {                                                  \
    if (!(cond)) return ::slx::fail(SLX_EINVAL, __VA_ARGS__); \
  } while (0)
