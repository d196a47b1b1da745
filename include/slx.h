/*
 * slx.h — C ABI of libslx.so, the gfx950 (MI355X) implementation of the
 * starlab-physics XPBD cloth step.
 *
 * The reference (medialab-ku/starlab-physics) is Python + Taichi: every entry
 * point below replaces one Taichi kernel (or a fixed sequence of them) and the
 * comment on each cites that kernel as file:line, paths relative to the
 * reference repository root.  The host side (Python, ctypes) keeps the
 * reference's Solver(...).forward(n_substeps) surface and passes raw device
 * pointers from PyTorch-ROCm tensors (tensor.data_ptr()) and the current HIP
 * stream (torch.cuda.current_stream().cuda_stream).
 *
 * Conventions
 *  - All pointers are device pointers unless the name says host.  Sizes are
 *    element counts.  Nothing here allocates device memory: buffers and
 *    workspaces belong to the caller (workspace sizes come from *_ws_bytes).
 *  - Every call is stream-ordered and asynchronous; no call synchronises the
 *    device except where stated.
 *  - Return value: SLX_OK (0) or a negative error; slx_last_error() gives the
 *    calling thread's last message.  Data-dependent conditions (candidate or
 *    stack overflow, out-of-range cells) are OR-ed into a caller-provided
 *    device int32 status word instead, so no call needs a hidden sync.
 *  - Vertex state layout (float4 per vertex, 16-byte aligned):
 *      x4 = (x, y, z, f*w)      position, f*w = fixed * m_inv
 *      y4 = (y, y, y, f*w)      predicted position (same w slot)
 *      v4 = (vx, vy, vz, f)     velocity, f = fixed (1 free, 0 pinned)
 *    plus fixed[nv] and m_inv[nv] as float arrays (meshTaichiWrapper.py:26-34).
 */
#ifndef SLX_H
#define SLX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SLX_ABI_VERSION 1

#define SLX_OK 0
#define SLX_EINVAL -1   /* bad argument (null pointer, negative size, ...) */
#define SLX_EHIP -2     /* HIP runtime error */
#define SLX_EWS -3      /* workspace too small */

#define SLX_STATUS_CAND_OVERFLOW 1   /* a candidate list exceeded its capacity */
#define SLX_STATUS_STACK_OVERFLOW 2  /* a traversal stack exceeded its depth */
#define SLX_STATUS_CELL_RANGE 4      /* a face centre fell outside the 32^3 cell grid */
#define SLX_STATUS_REC_OVERFLOW 8    /* a vertex received more than rec_cap contact deltas (dropped) */
#define SLX_STATUS_PAIR_OVERFLOW 16  /* a query held more than pair_cap contacting pairs (dropped) */

typedef void* slx_stream_t; /* hipStream_t */

int slx_abi_version(void);
const char* slx_last_error(void);

/* Kernel timer (no reference counterpart: measurement support).  When enabled,
 * every kernel launch records a hipEvent pair around itself on its own stream,
 * up to max_launches launches; slx_timer_read sums the elapsed time of the
 * launches of one kernel name (it synchronises on those events).
 * max_launches = 0 disables the timer. */
int slx_timer_enable(int max_launches);
int slx_timer_read(const char* name, double* total_ms, int* count);
/* Time only the launches named `name` (null or "": all), so that the timed
 * region carries one event pair per launch of that kernel and nothing else. */
int slx_timer_only(const char* name);
int slx_timer_names(char* buf, int buf_len);

/* ---------------- XPBD Jacobi solve (framework/physics/XPBD.py) ---------------- */

/* compute_y, XPBD.py:115-118: y = x + f*v*dt + g*dt*dt; y4.w = x4.w. */
int slx_compute_y(int nv, const float* x4, const float* v4, float* y4, float dt, float gx,
                  float gy, float gz, slx_stream_t stream);

/* init_variables + solve_spring_constraints_x + update_dx, XPBD.py:345-348,
 * 140-184, 383-387, as a per-vertex gather over the SELL-64 adjacency
 * (slice_off[nslices+1], slice_width[nslices], entries[2*total]): each vertex sums
 * its incident constraints' deltas in ascending constraint order (stretch on y,
 * bending on x), then y_out = y_in + f*(dx/nc).  k_s / k_b are the stretch and
 * bending factors stiffness*dt^2 (XPBD.py:354-357). */
int slx_spring_pass(int nv, const float* x4, const float* fixed, const float* y4_in,
                    float* y4_out, const int32_t* slice_off, const int32_t* slice_width,
                    const int32_t* entries, float k_s, float k_b, slx_stream_t stream);

/* compute_velocity then update_x, XPBD.py:340-343, 332-337:
 * v = (1-damping)*f*(y-x)/dt; x += dt*v.  In place on x4/v4. */
int slx_update_vx(int nv, float* x4, float* v4, const float* y4, const float* fixed,
                  float damping, float dt, slx_stream_t stream);

/* n_substeps of the collision-free step (XPBD.py:460-469 with
 * enable_collision_handling False, n_iters spring passes per substep) with
 * predict fused into the first pass and velocity/position update fused into the
 * last.  x4/v4 hold the state on entry and on return; x4_alt/v4_alt/y4a/y4b are
 * scratch of the same size. */
int slx_substeps_plain(int nv, float* x4, float* v4, float* x4_alt, float* v4_alt, float* y4a,
                       float* y4b, const float* fixed, const int32_t* slice_off,
                       const int32_t* slice_width, const int32_t* entries, int n_substeps,
                       int n_iters, float dt, float gx, float gy, float gz, float k_s, float k_b,
                       float damping, slx_stream_t stream);

/* Launch shape of every spring-pass kernel (no reference counterpart: set by the
 * autotuner): threads per block {64, 128, 256} x vertices per thread {1, 2, 4}.
 * The library default (a frame descriptor may override it for its call);
 * results do not depend on it. */
int slx_set_spring_launch(int block, int vpt);

/* (f, m_inv) -> x4.w = y4.w = f*m_inv, v4.w = f (set_fixed_vertices,
 * XPBD.py:389-395, with the per-vertex products the kernels read). */
int slx_set_fixed(int nv, const float* fixed, const float* m_inv, float* x4, float* v4,
                  float* y4, slx_stream_t stream);

/* ---------------- device primitives ---------------- */

/* Exclusive scan of int32 (in -> out), total written to *total_dev when non-null.
 * n_dev (nullable) holds the element count on the device (<= n_max). */
size_t slx_scan_ws_bytes(int n_max);
int slx_exclusive_scan(int n_max, const int32_t* n_dev, const int32_t* in, int32_t* out,
                       int32_t* total_dev, void* ws, size_t ws_bytes, slx_stream_t stream);

/* Stable LSD radix sort of (keys, vals) in place by the low key_bits bits
 * (8-bit digits; the reference sorts with 6-bit digits, lbvh_cell.py:524-562, a
 * stable sort by the same key gives the same order).  n_dev as above. */
size_t slx_radix_sort_ws_bytes(int n_max);
int slx_radix_sort_pairs(int n_max, const int32_t* n_dev, uint32_t* keys, int32_t* vals,
                         int key_bits, void* ws, size_t ws_bytes, slx_stream_t stream);

/* slx_radix_sort_pairs without a workspace argument and without any allocation
 * (no hipMalloc, hipFree or allocating library call).  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Its scratch
 * is a static device array of 18.9 MB per GPU, held until the process ends,
 * which limits n_max to 2,097,152 pairs (larger n_max: SLX_EINVAL with a
 * message).  Every call on one GPU shares that GPU's scratch, so calls that
 * overlap on one GPU (on different streams, or from different threads) race;
 * calls in one stream's order do not. */
int slx_radix_sort_pairs_noalloc(int n_max, const int32_t* n_dev, uint32_t* keys, int32_t* vals,
                                 int key_bits, slx_stream_t stream);

/* ---------------- broad phase (framework/collision/lbvh_cell.py) ---------------- */

/* computeAABB_faces, meshTaichiWrapper.py:350-361: face boxes of x4 -+ padding,
 * float4 per face (w = face id). */
int slx_face_aabbs(int nf, const int32_t* faces, const float* x4, float padding, float* fmin4,
                   float* fmax4, slx_stream_t stream);

/* computeAABB, meshTaichiWrapper.py:334-347: box6 = (min - padding, max + padding),
 * left on the device (the reference returns it to the host). */
size_t slx_mesh_aabb_ws_bytes(void);
int slx_mesh_aabb(int nv, const float* x4, float padding, float* box6, void* ws,
                  size_t ws_bytes, slx_stream_t stream);

/* LBVH_CELL.__init__, lbvh_cell.py:17-80: Morton codes of the 32^3 cell centres
 * of [0,1e3]^3 (assign_cell_morton 122-152), stable radix sort (553-562), Karras
 * 2012 tree (407-430).  codes/sorted_ids [32768], child_a/child_b/range_l/range_r
 * [32767], parent [65535]. */
size_t slx_cell_topology_ws_bytes(void);
int slx_cell_topology(int32_t* codes, int32_t* sorted_ids, int32_t* child_a, int32_t* child_b,
                      int32_t* parent, int32_t* range_l, int32_t* range_r, int32_t* sched, void* ws,
                      size_t ws_bytes, slx_stream_t stream);

/* Refit schedule of a fixed topology over n_leaves leaves (sched holds
 * slx_refit_sched_ints(n_leaves) int32; the topology entry points above and below
 * write it).  The reference refits with one thread per leaf and global arrival
 * counters (lbvh_cell.py:490-522); the schedule instead cuts the tree into subtrees
 * of at most sub_leaves leaves, each refit by one workgroup in LDS, and the nodes
 * above them: their maximal subtrees of at most 64 x sub_leaves leaves one
 * workgroup each, then the rest in one workgroup, so no device-wide fence is
 * needed.  The boxes are the same.  slx_set_refit_shape sets sub_leaves (1..2048,
 * default 1024) and how many top nodes of one workgroup keep their state in LDS
 * (0..2048, default 2048; at most 512 are kept, the others count arrivals in
 * global memory) for schedules made afterwards. */
size_t slx_refit_sched_ints(int n_leaves);
int slx_refit_schedule(int n_leaves, const int32_t* child_b, const int32_t* parent,
                       const int32_t* range_l, const int32_t* range_r, int32_t* sched,
                       slx_stream_t stream);
int slx_set_refit_shape(int sub_leaves, int top_lds);

/* LBVH_CELL.build, lbvh_cell.py:637-653: faces bucketed by cell (stable, ascending
 * face id inside a cell), inclusive cell prefix, leaf slot i = row-major cell i
 * (box = cell centre U its faces' boxes), bottom-up refit (490-522).  Node arrays
 * hold 65535 nodes: num [65535], nlo4/nhi4 float4 [65535]; leaf_l/leaf_r/prefix
 * [32768]; face_cell/sorted_faces [nf]. */
size_t slx_cell_build_ws_bytes(int nf);
int slx_cell_build(int nf, const float* fmin4, const float* fmax4, const float* box6,
                   const int32_t* child_a, const int32_t* child_b, const int32_t* parent,
                   const int32_t* sched, int32_t* face_cell, int32_t* prefix, int32_t* sorted_faces, int32_t* num,
                   int32_t* leaf_l, int32_t* leaf_r, float* nlo4, float* nhi4, int32_t* status,
                   void* ws, size_t ws_bytes, slx_stream_t stream);

/* Per-triangle LBVH (the reference's unused per-face path, lbvh_cell.py:153-194,
 * 237-247, 383-405, completed): 30-bit Morton codes of face-box centres in the
 * cube [lo, lo + e] of box6 (e = its largest extent, one scale on every axis, so
 * a thin sheet keeps its bits for its width), stable radix sort of (code, face
 * id), Karras tree with an index tie-break for
 * equal codes, refit.  codes/sorted_faces [nf], child_a/child_b [nf-1], parent
 * [2nf-1], nlo4/nhi4 float4 [2nf-1]. */
size_t slx_tri_build_ws_bytes(int nf);
int slx_tri_build(int nf, const float* fmin4, const float* fmax4, const float* box6,
                  int32_t* codes, int32_t* sorted_faces, int32_t* child_a, int32_t* child_b,
                  int32_t* parent, int32_t* range_l, int32_t* range_r, int32_t* sched,
                  float* nlo4, float* nhi4, float* quad4, void* ws, size_t ws_bytes,
                  slx_stream_t stream);
/* quad4 (nullable): the triangle tree's 4-wide node array, float4 [8 (nf-1)], kept
 * by every build and refit: internal node p holds its four grandchildren's boxes
 * (left-left, left-right, right-left, right-right; a leaf child in the first slot
 * of its pair with the other slot empty) as (lo, node) (hi, face or -1).  The
 * traversal then tests two levels per load; candidate lists are unchanged. */
/* Refit of the triangle LBVH (topology and schedule kept): leaf k gets the padded
 * box of face sorted_faces[k] computed from x4 (the face boxes of
 * computeAABB_faces, meshTaichiWrapper.py:350-361, fused in; fmin4/fmax4 are not
 * rewritten), internal nodes the union of their children. */
int slx_tri_refit(int nf, const int32_t* faces, const float* x4, float padding,
                  const int32_t* sorted_faces, const int32_t* child_a, const int32_t* child_b,
                  const int32_t* parent, const int32_t* sched, float* nlo4, float* nhi4,
                  float* quad4, slx_stream_t stream);

/* traverse_cell_bvh_single, lbvh_cell.py:655-714, for nq query points p -+ padding
 * (p = xyz of q4).  mode 0 = cell tree (queries whose centre cell id exceeds 32767
 * are skipped, as in the reference), 1 = triangle tree.  Candidates in traversal
 * order to cand[j*nq + q], j < cap; cand_n[q] = min(count, cap); overflow sets a
 * status bit (the reference writes past its 40-entry cache). */
int slx_traverse(int mode, int nq, const float* q4, float padding, int n_leaves,
                 const int32_t* num, const int32_t* sorted_faces, const float* nlo4,
                 const float* nhi4, const float* quad4, const float* fmin4, const float* fmax4,
                 const float* box6, int cap, int32_t* cand, int32_t* cand_n, int32_t* status,
                 slx_stream_t stream);

/* The edge broad phase of XPBD.py:217-232 (commented out in the reference): the
 * same traversal for ne edge boxes, min/max of x4[edges[2e]], x4[edges[2e+1]]
 * -+ padding.  Arguments otherwise as slx_traverse (q = edge).  A dynamic tree is
 * queried with its own grid (the reference's commented query of lbvh_dy passes
 * the static tree's cell size and origin).  order (nullable, [ne]): the order in
 * which edges are walked; redo (nullable, [1 + ne], triangle trees): the two-tier
 * walk (slx_set_walk_stack entries first, the edges that overflow it listed in
 * redo and walked again with the full stack; redo[0] = their count).  The lists
 * do not depend on either. */
int slx_traverse_edges(int mode, int ne, const int32_t* edges, const float* x4, float padding,
                       int n_leaves, const int32_t* num, const int32_t* sorted_faces,
                       const float* nlo4, const float* nhi4, const float* quad4, const float* fmin4,
                       const float* fmax4, const float* box6, int cap, int32_t* cand,
                       int32_t* cand_n, int32_t* status, const int32_t* order, int32_t* redo,
                       slx_stream_t stream);

/* Batched instances (no reference counterpart beyond the mesh concatenation of
 * framework/meshio/concat.py:11-58): instances packed CSR-style (vert_base,
 * face_base [n_inst+1], face_inst [nf]); per-instance boxes box6[6*n_inst];
 * Morton codes in each instance's box; stable sort by code then by instance
 * (one segmented sort); one Karras tree over the composite keys
 * (instance, code, index) in which roots[i] is the subtree holding exactly
 * instance i's faces; refit.  Includes the face boxes (x4, padding). */
size_t slx_seg_build_ws_bytes(int nf, int n_inst);
int slx_seg_build(int nv, int nf, int n_inst, const int32_t* faces, const float* x4, float padding,
                  const int32_t* vert_base, const int32_t* face_base, const int32_t* face_inst,
                  float* fmin4, float* fmax4, float* box6, int32_t* codes, int32_t* sorted_faces,
                  int32_t* sorted_inst, int32_t* child_a, int32_t* child_b, int32_t* parent,
                  int32_t* range_l, int32_t* range_r, int32_t* roots, int32_t* sched, float* nlo4,
                  float* nhi4, float* quad4, void* ws, size_t ws_bytes, slx_stream_t stream);
/* Candidate cache of one query set against a triangle tree (no reference
 * counterpart; the lists it gives are the traversal's, bit for bit).  Within a
 * frame the tree keeps its topology and its walks emit leaves in one fixed
 * order, so a walk with the query box enlarged by `margin` can store each
 * query's leaves, and later passes filter them by the exact box instead of
 * walking: exact while the query moved d and the tree's vertices D since the
 * cache was built with d + D below the margin (less a rounding allowance).
 * Queries the cache cannot serve are walked in full. */
typedef struct slx_cand_cache {
  int32_t cap;                           /* leaves stored per query */
  float margin;                          /* enlargement of the query box (length) */
  int32_t* leaves;                       /* leaf node ids: int4 [(cap + 3) / 4][nq] (16-byte aligned,
                                            4 ceil(cap/4) nq int32), four per processing slot and row */
  int32_t* count;                        /* [nq] leaves found (> cap or INT_MAX: incomplete) */
  float* q0;                             /* float4 [nq] query positions at the build */
  int32_t* redo;                         /* [nq + 1] queries walked in full ([0] = how many) */
} slx_cand_cache;

/* Exact candidate lists of nq queries against a triangle tree (as slx_traverse
 * mode 1) in two tiers: a first walk with a shallow stack (more waves in flight)
 * and, for the queries it cannot serve, a walk with the full stack.  kind 0 walks
 * (only cache->redo is used); kind 1 walks with the enlarged box and fills the
 * cache (call it right after the tree's snapshot, below); kind 2 filters the
 * cache, disp (device, nullable = 0) being the tree's vertex displacement since
 * the cache was filled.  order (nullable): processing order of the queries, the
 * same for every call on one cache (its entries are indexed by that order). */
int slx_traverse_cached(int kind, int nq, const float* q4, float padding, int n_leaves,
                        const float* nlo4, const float* nhi4, const float* quad4,
                        const slx_cand_cache* cache, const uint32_t* disp, int cap, int32_t* cand,
                        int32_t* cand_n, int32_t* status, const int32_t* order, slx_stream_t stream);
/* Stack depth of the first-tier walk: 16, 24 (default), 32 or 64 entries. */
int slx_set_walk_stack(int depth);
/* Workgroups (1..65535, default 2048) of a walk whose query count is on the
 * device (the second tier, the edge gate's live lists): such a walk loops over
 * its queries, so one with few or none ends at once.  Results do not depend on
 * it (a tuning and test knob, no counterpart in the reference). */
int slx_set_active_grid(int blocks);
/* x_snap = x (float4 [nv]); disp = 0. */
int slx_tree_snapshot(int nv, const float* x4, float* x_snap, uint32_t* disp, slx_stream_t stream);
/* disp = max over vertices of |x - x_snap|_inf (float bits; infinity for NaN). */
int slx_tree_displacement(int nv, const float* x4, const float* x_snap, uint32_t* disp,
                          slx_stream_t stream);

/* Traversal of a batched tree: query q starts at roots[q_inst[q]]. */
int slx_traverse_seg(int nq, const float* q4, const int32_t* q_inst, const int32_t* roots,
                     float padding, int nf, const float* nlo4, const float* nhi4, const float* quad4,
                     int cap, int32_t* cand, int32_t* cand_n, int32_t* status, slx_stream_t stream);

/* ---------------- narrow phase (collision_constraints_x.py, distance.py) ---------------- */

/* solve_collision_constraints_x + update_dx, XPBD.py:236-264, 383-387, with
 * __vt_st / __vt_dy / __tv_st (collision_constraints_x.py:5-310): the three
 * candidate sets (any may be null) in the reference's loop order; deltas are
 * accumulated per vertex in that serial order without float atomics; y4 updated
 * in place.  pair_num [2nv+nvs] (nullable) receives the recorded pair counts.
 * rec_cap = contact deltas kept per dynamic vertex and pass from the second and
 * third sets (more -> dropped, status bit); (nv+nvs)*cap*4 must stay below 2^32.
 *
 * Pair records for the velocity pass (the reference's vt_st_pair / vt_dy_pair /
 * tv_st_pair fields with their g and schur, XPBD.py:51-73): with pair_rec and
 * pair_cnt non-null, each query q of the 2nv+nvs keeps the pairs among its first
 * cap records that are in contact (d < dhat; the others have zero gradients and
 * no effect on velocities), at most pair_cap (more -> dropped, status bit
 * SLX_STATUS_PAIR_OVERFLOW; pair_cap = cap never drops).  Layout: float4
 * pair_rec[(j*4 + w) * (2nv+nvs) + q] for pair j < pair_cnt[q], w = 0..3:
 * (g0, schur), (g1, face bits), (g2, type bits), (g3, 0). */
size_t slx_collision_ws_bytes(int nv, int nvs, int rec_cap);
int slx_collision_pass(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st,
                       float* y4, const float* xs4, const float* fixed, const float* m_inv,
                       float dhat, float k_col, int cap, const int32_t* c_vt_st,
                       const int32_t* n_vt_st, const int32_t* c_vt_dy, const int32_t* n_vt_dy,
                       const int32_t* c_tv_st, const int32_t* n_tv_st, int32_t* pair_num,
                       int pair_cap, int32_t* pair_cnt, float* pair_rec,
                       int rec_cap, int32_t* status, void* ws, size_t ws_bytes,
                       slx_stream_t stream);

/* Edge-edge sets of one collision pass (the reference's ee_st / ee_dy fields,
 * XPBD.py:75-89, and loops, 266-281 and 313-330, all commented out there).
 * Query e (one per dynamic edge) tests the three sides (face_edge_indices) of
 * each candidate face of the static tree (set 0, __ee_st) and then of the
 * dynamic tree (set 1, __ee_dy, skipping sides that share a vertex with e).
 * Candidates column-major [cap][ne] with counts, as slx_traverse_edges writes
 * them; a null candidate array leaves its set out.  Pair records (the
 * reference's ee_*_pair fields): pair_num[s*ne + e] (nullable) counts the
 * contacting pairs (d < dhat) of set s, uncapped; with pair_rec and pair_cnt
 * non-null the first pair_cap of them are stored (more -> dropped, status bit
 * SLX_STATUS_PAIR_OVERFLOW), float4 pair_rec[(j*4 + w) * 2ne + s*ne + e] for
 * j < pair_cnt[s*ne + e]: (g0, schur), (g1, other edge bits), (g2, type bits),
 * (g3, 0).  18*ne*cap and 6*ne*pair_cap must stay below 2^32. */
typedef struct slx_edge_sets {
  int32_t ne;
  const int32_t* edges_dy;  /* [ne][2] */
  const int32_t* edges_st;  /* [nes][2] */
  const int32_t* fe_dy;     /* [F][3] edge ids of each face's sides */
  const int32_t* fe_st;     /* [Fs][3] */
  int32_t cap;
  const int32_t* c_st;
  const int32_t* n_st;
  const int32_t* c_dy;
  const int32_t* n_dy;
  int32_t pair_cap;
  int32_t* pair_num;        /* [2 ne] */
  int32_t* pair_cnt;        /* [2 ne] */
  float* pair_rec;          /* float4 [pair_cap][4][2 ne] */
  const int32_t* order;     /* nullable: [ne] order in which slx_frame walks the dynamic tree for the
                               edges (a spatially coherent one lets a wavefront's queries share tree
                               nodes); results do not depend on it */
  int32_t* redo;            /* nullable: [2 (1 + ne)] lists of the edges a shallow-stack walk could
                               not serve (static tree, dynamic tree), for slx_frame's two-tier
                               edge walk of triangle trees */
  /* the edge gate (slx_frame only, triangle trees; both nullable = off): type 8
     needs (u x v).w >= 1e-4, and |(u x v).w| <= |u| |v| |w| is bounded per edge
     from the longest edges and the largest y - x displacement of the pass; an
     edge whose bound, with rounding allowances, stays below 1e-4 is neither
     walked nor tested (its candidate list is not written; it could not have
     acted).  The dynamic set is gated in substeps whose dynamic tree was refit
     from the current x.  slx_collision_pass_ee ignores these fields. */
  uint32_t* gate;           /* [8] scratch */
  int32_t* live;            /* [2 (1 + ne)] lists of the edges the gate keeps (static, dynamic) */
  int32_t nes;              /* static edges (the rows of edges_st) */
} slx_edge_sets;

/* slx_collision_pass with the edge-edge constraints after the vertex-triangle
 * ones (collision_constraints_x.py:312-507: type 8 only, edges further than 1e-3
 * from parallel, d = d_EE < dhat, ld = k (dhat - d) / (k schur + 1) on the
 * dynamic points), all accumulated per vertex in the reference's serial order
 * before update_dx.  ee null: slx_collision_pass. */
int slx_collision_pass_ee(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st,
                          float* y4, const float* xs4, const float* fixed, const float* m_inv,
                          float dhat, float k_col, int cap, const int32_t* c_vt_st,
                          const int32_t* n_vt_st, const int32_t* c_vt_dy, const int32_t* n_vt_dy,
                          const int32_t* c_tv_st, const int32_t* n_tv_st, int32_t* pair_num,
                          int pair_cap, int32_t* pair_cnt, float* pair_rec, const slx_edge_sets* ee,
                          int rec_cap, int32_t* status, void* ws, size_t ws_bytes,
                          slx_stream_t stream);

/* ---------------- sewing constraints (XPBD.py:405-441) ---------------- */

/* solve_sewing_constraints_x as a Jacobi pass of its own (zero dx/nc, the
 * pairs, update_dx; the reference's commented-out call accumulates after its
 * last update_dx, where nothing would apply it): pair (v1, v2) pulls its points
 * together with ld = k C / (k (f1 w1 + f2 w2) + 1), C = |y2 - y1|.  pairs int2
 * [np]; the n_sewn distinct vertices in sew_verts, each with its pair entries
 * (2 pair + side, in pair order) in sew_list[sew_off[i] .. sew_off[i+1]).  y4.w
 * must hold f*m_inv.  ws: float4 [n_sewn].  A coincident pair adds no delta. */
int slx_sewing_pass(int n_sewn, const int32_t* sew_verts, const int32_t* sew_off, const int32_t* sew_list,
                    const int32_t* pairs, float* y4, const float* fixed, float k, float* ws,
                    slx_stream_t stream);

/* ---------------- velocity pass (collision_constraints_v.py) ---------------- */

/* compute_velocity, XPBD.py:340-343: v = (1-damping)*f*(y-x)/dt; v4.w = f. */
int slx_compute_velocity(int nv, const float* x4, float* v4, const float* y4, const float* fixed,
                         float damping, float dt, slx_stream_t stream);
/* update_x, XPBD.py:332-337: x += dt*v. */
int slx_update_x(int nv, float* x4, const float* v4, float dt, slx_stream_t stream);
/* solve_constraints_v, XPBD.py:374-381: zero dv/nc; with pair records (see
 * slx_collision_pass) the velocity-level contact and friction constraints of
 * each query set (solve_collision_constraints_v, XPBD.py:283-311, with __vt_st /
 * __vt_dy / __tv_st of collision_constraints_v.py:5-570) accumulated per vertex in
 * the reference's serial order; then update_dv (XPBD.py:397-404): dv /= nc where
 * nc > 0, v += f*dv where m_inv > 0.  v4 in place.  pair_cnt null: update_dv
 * alone (the reference without collision handling).
 *   vs4      static vertex velocities (float4, nullable = at rest, as the
 *            reference's static meshes);
 *   m_inv_st static inverse masses, read by __vt_st type 3 at the dynamic vertex's
 *            index (collision_constraints_v.py:64), 0 past nvs or when null;
 *   x4       nullable: when given, update_x (x += dt*v) is fused in.
 * Workspace: slx_collision_ws_bytes(nv, nvs, rec_cap); (nv+nvs)*pair_cap*8 must
 * stay below 2^32. */
int slx_velocity_pass(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st, float* v4,
                      const float* vs4, const float* fixed, const float* m_inv, const float* m_inv_st,
                      float mu, int pair_cap, const int32_t* pair_cnt, const float* pair_rec, float* x4,
                      float dt, int rec_cap, int32_t* status, void* ws, size_t ws_bytes,
                      slx_stream_t stream);
/* slx_velocity_pass with the edge-edge pairs recorded by slx_collision_pass_ee
 * after the vertex-triangle ones (collision_constraints_v.py:573-878, normal
 * only: __ee_st with Cv = g0.v0 + g1.v1 + g3.vs3 -- no g2 term, :782 -- moves the
 * dynamic edge's points by f w (-Cv/schur) g; __ee_dy with dvn = sum g.v moves all
 * four by -f w g (dvn/schur); both only when closing).  ee null:
 * slx_velocity_pass. */
int slx_velocity_pass_ee(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st, float* v4,
                         const float* vs4, const float* fixed, const float* m_inv,
                         const float* m_inv_st, float mu, int pair_cap, const int32_t* pair_cnt,
                         const float* pair_rec, const slx_edge_sets* ee, float* x4, float dt,
                         int rec_cap, int32_t* status, void* ws, size_t ws_bytes,
                         slx_stream_t stream);

/* ---------------- graph-coloured Gauss-Seidel ---------------- */

/* Greedy colouring of the stretch edges (HOST pointers; graph_coloring.py:131-149):
 * edges in index order, each takes the lowest colour not used by an already-
 * coloured edge sharing a vertex; order[] lists edges colour by colour, ascending
 * id inside a colour (stable counting sort, graph_coloring.py:172-201);
 * offsets[n_colors+1].  Returns n_colors, or a negative error. */
int slx_greedy_edge_coloring(int nv, int ne, const int32_t* edges, int max_colors,
                             int32_t* color, int32_t* order, int32_t* offsets);

/* One Gauss-Seidel iteration in place on y4: the bending constraints (on x) as a
 * Jacobi-averaged pass, then the stretch edges colour by colour (gs_edges: int4
 * (v0, v1, l0 bits, 0) in colour order; color_off_host: HOST [n_colors+1]).
 * The reference colours edges but has no sweep; this sweep is the package's. */
int slx_gs_pass(int nv, const float* x4, const float* fixed, float* y4, const int32_t* slice_off,
                const int32_t* slice_width, const int32_t* entries, const int32_t* gs_edges,
                const int32_t* color_off_host, int n_colors, float k_s, float k_b,
                slx_stream_t stream);

/* ---------------- learned warm start (no reference counterpart) ---------------- */

/* Number of FP32 parameters of the per-vertex MLP (24 -> 64 -> 64 -> 64 -> 3,
 * GELU; layout W1[64][32] b1 W2[64][64] b2 W3[64][64] b3 W4[32][64] b4[32],
 * weights [out][in], pads zero). */
int slx_mlp_num_params(void);
/* 24 features per vertex (FP32, feat[nv][24]) from x4, v4 and the predicted y4
 * over the SELL adjacency; inv_s = 1 / mean rest length. */
int slx_features(int nv, const float* x4, const float* v4, const float* y4, const int32_t* slice_off,
                 const int32_t* slice_width, const int32_t* entries, float dt, float inv_s,
                 float* feat, slx_stream_t stream);
/* Warm start of a substep's Jacobi iterate: features of y4_pred -> MLP on BF16
 * MFMA (w_bf16 weights, w_f32 biases, feature mean / inverse std) ->
 * y4_out = y4_pred + fixed * out * out_scale. */
int slx_warm_start(int nv, const float* x4, const float* v4, const float* y4_pred, float* y4_out,
                   const float* fixed, const int32_t* slice_off, const int32_t* slice_width,
                   const int32_t* entries, float dt, float inv_s, const void* w_bf16,
                   const float* w_f32, const float* fmean, const float* fistd, float out_scale,
                   slx_stream_t stream);
/* Gradient of mean_n sum_3 (out - target)^2 w.r.t. the parameters (forward and
 * backward on BF16 MFMA, FP32 accumulation): per-block partials part[n_blocks][P],
 * loss partials loss_part[n_blocks], and (nullable) their sum in block order. */
int slx_mlp_grad(int n, const float* feat, const float* target, const void* w_bf16,
                 const float* w_f32, const float* fmean, const float* fistd, int n_blocks,
                 float* part, float* loss_part, float* grad, slx_stream_t stream);
/* Adam step on the FP32 master weights from the summed partials; refreshes the
 * BF16 copy. step counts from 1. */
int slx_mlp_adam(int n_blocks, const float* part, float* w_f32, void* w_bf16, float* m1, float* m2,
                 float lr, float beta1, float beta2, float eps, int step, slx_stream_t stream);
int slx_mlp_to_bf16(const float* w_f32, void* w_bf16, slx_stream_t stream);
/* *out = max(*out, max_v |a - b|_inf) as float bits (caller zeroes *out): the
 * Jacobi update residual. */
int slx_max_update(int nv, const float* a4, const float* b4, uint32_t* out, slx_stream_t stream);

/* Residual solve (no reference counterpart: the reference runs one pass per
 * substep, XPBD.py:443-469, and repeating it does not converge because its bending
 * term acts on the unchanging x).  Anchored at the reference pass's result y1, the
 * substep's implicit-Euler stretch problem
 *   min sum_v |y_v - y1_v|^2/(2 w_v) + sum_e kappa C_e(y)^2/2
 * (w = fixed*m_inv, kappa = stiffness*dt^2, C_e the stretch constraints of
 * XPBD.py:146-167) is solved by Jacobi with the diagonal Gauss-Newton step
 * y_v -= f_v (y_v - y1_v + w_v sum kappa C n) / (1 + w_v sum kappa); the residual of
 * an iterate is its largest step component.  Pass k reads y_k (ya for even k, yb
 * for odd) and writes y_{k+1} into the other buffer, and max-reduces the residual
 * of y_k into res[k] (float bits).  A pass whose predecessor's residual is <= tol,
 * or that follows such a pass, only sets *conv = min(*conv, k-1) and returns, so
 * the first k with res[k] <= tol names the answer y_k.  slx_solve_reset zeroes
 * res[0..n_res) and sets *conv = INT_MAX (no memset; graph-safe). */
int slx_solve_reset(int n_res, uint32_t* res, int32_t* conv, slx_stream_t stream);
int slx_implicit_passes(int nv, const float* fixed, const float* y_anchor, float* ya, float* yb,
                        const int32_t* slice_off, const int32_t* slice_width, const int32_t* entries,
                        float kappa, int k0, int count, float tol, uint32_t* res, int32_t* conv,
                        slx_stream_t stream);

/* ---------------- one frame (Solver.forward, XPBD.py:443-469) ---------------- */

typedef struct slx_tree {
  int32_t mode;              /* 0 = cell LBVH, 1 = triangle LBVH */
  int32_t nf;                /* faces */
  float* fmin4;              /* float4 [nf] face boxes */
  float* fmax4;
  float* box6;               /* mesh box (after padding) */
  int32_t* codes;            /* triangle: Morton codes [nf] (sorted) */
  int32_t* sorted_faces;     /* [nf] */
  int32_t* child_a;          /* cell: static topology [32767]; triangle: [nf-1] */
  int32_t* child_b;
  int32_t* parent;           /* cell: [65535]; triangle: [2nf-1] */
  int32_t* num;              /* cell: faces under each node [65535] */
  int32_t* face_cell;        /* cell: [nf] */
  int32_t* prefix;           /* cell: inclusive prefix [32768] */
  int32_t* leaf_l;           /* cell: [32768] */
  int32_t* leaf_r;
  float* nlo4;               /* float4 [nodes] */
  float* nhi4;
  /* batched instances (triangle mode with n_inst > 1, see slx_seg_build) */
  int32_t n_inst;            /* 0 or 1: a single mesh */
  const int32_t* vert_base;  /* [n_inst+1] */
  const int32_t* face_base;  /* [n_inst+1] */
  const int32_t* face_inst;  /* [nf] */
  const int32_t* vert_inst;  /* [nv] instance of each vertex of this tree's mesh (its queries) */
  int32_t* sorted_inst;      /* [nf] */
  int32_t* range_l;          /* triangle: leaf range of each internal node [nf-1] */
  int32_t* range_r;
  int32_t* roots;            /* [n_inst] */
  int32_t* sched;            /* refit schedule, slx_refit_sched_ints(leaves) (cell: of the static topology) */
  const int32_t* vert_order; /* [nv] order in which this tree's mesh's vertices are processed as queries
                                (nullable; e.g. Morton order, so that a wavefront's queries are neighbours) */
  float* quad4;              /* triangle: 4-wide node array, float4 [8 (nf-1)] (nullable, see below) */
  /* candidate caches over this tree (slx_frame): its vertices at the build and
     their displacement since (float bits); nullable */
  float* x_snap;             /* float4 [nv] */
  uint32_t* disp;            /* [1] */
} slx_tree;

typedef struct slx_frame_ctx slx_frame_ctx;   /* opaque */

typedef struct slx_frame_desc {
  /* dynamic mesh */
  int32_t nv;
  float* x4;
  float* v4;
  float* y4;
  float* y4_alt;
  const float* fixed;
  const float* m_inv;
  const int32_t* faces;                  /* [3 nf] */
  const int32_t* adj_slice_off;
  const int32_t* adj_slice_width;
  const int32_t* adj_entries;
  /* static mesh (nvs = 0 when absent) */
  int32_t nvs;
  const float* xs4;
  const int32_t* faces_st;
  /* per-substep scalars (float32, as the reference passes them) */
  float dt, gx, gy, gz, k_s, k_b, damping;
  float padding, dhat, k_col;
  int32_t n_substeps, n_iters;
  int32_t collision;                     /* enable_collision_handling */
  int32_t refit_every_substep;           /* triangle tree: refit from x at each substep */
  int32_t rebuild_static;                /* rebuild the static tree this frame */
  int32_t cap, rec_cap;                  /* candidates per query; contact deltas per vertex */
  int32_t method;                        /* 0 = Jacobi, 1 = graph-coloured Gauss-Seidel */
  const int32_t* gs_edges;               /* GS: int4 (v0, v1, l0 bits, 0) per stretch edge, colour order */
  const int32_t* gs_color_off;           /* GS: HOST array [n_colors+1] of colour starts */
  int32_t n_colors;
  slx_tree dyn, st;
  int32_t* cand;                         /* [cap*nv] vt_st, [cap*nv] vt_dy, [cap*nvs] tv_st */
  int32_t* cand_n;                       /* [2nv + nvs] */
  int32_t* pair_num;                     /* [2nv + nvs], nullable */
  int32_t* status;                       /* device status word */
  void* ws;
  size_t ws_bytes;
  /* velocity pass (enable_velocity_update, XPBD.py:29, 467-468) */
  int32_t velocity;
  float mu;
  const float* vs4;                      /* static vertex velocities, nullable */
  const float* m_inv_st;                 /* static inverse masses, nullable */
  int32_t pair_cap;                      /* contacting pairs kept per query */
  int32_t* pair_cnt;                     /* [2nv + nvs] */
  float* pair_rec;                       /* float4 [4 pair_cap (2nv + nvs)] */
  /* candidate caches of the query sets vt_st, vt_dy, tv_st (triangle trees, single
     or batched -- a batched tree's queries walk from their instance's root; cap 0 =
     walk every pass).  Rebuilt at the first query of each
     set in every cache_every-th substep (<= 0: once per frame); the dynamic tree
     needs x_snap and disp. */
  slx_cand_cache cache[3];
  int32_t cache_every;
  /* sewing constraints (XPBD.py:418-441), after the collision pass of each
     iteration (n_sewn = 0: none); k_sew = sewing stiffness * dt^2 */
  int32_t n_sewn;
  const int32_t* sew_verts;
  const int32_t* sew_off;
  const int32_t* sew_list;
  const int32_t* sew_pairs;
  float k_sew;
  float* sew_ws;
  /* edge-edge constraints (XPBD.py:217-232, 266-281, 313-330, commented out
     there; ee.ne = 0: off).  Each collision pass queries the dynamic edges'
     boxes against the static tree (when there is a static mesh) and the dynamic
     tree, into the caller's buffers ee.c_st / ee.n_st / ee.c_dy / ee.n_dy
     ([ee.cap][ne] and [ne]; written by the frame), and applies the edge-edge
     constraints after the vertex-triangle ones; the pass that records the
     vertex-triangle pairs for the velocity pass records the edge pairs too
     (ee.pair_cnt / ee.pair_rec), and the velocity pass applies them.  Single
     meshes only (no batched trees). */
  slx_edge_sets ee;
  /* per-solver execution state (no reference counterpart).  ctx: the side
     streams and events this frame forks onto (slx_frame_ctx_create); null uses
     one shared set per device, so two solvers that leave it null must not run
     frames concurrently.  Launch knobs for this call only, 0 = the library
     default (slx_set_spring_launch / slx_set_walk_stack / slx_set_active_grid /
     slx_set_refit_shape); results never depend on them. */
  slx_frame_ctx* ctx;
  int32_t spring_block;                  /* 64, 128 or 256 */
  int32_t spring_vpt;                    /* 1, 2 or 4 */
  int32_t walk_stack;                    /* 16, 24, 32 or 64 */
  int32_t active_grid;                   /* 1..65535 */
  int32_t refit_sub;                     /* 1..2048 */
  int32_t refit_top;                     /* 1..2048 */
  int32_t walk_wave;                     /* cloth-vertex walks of the cloth tree: 1 = one shared walk per
                                            wavefront over the 4-wide nodes, -1 = one walk per lane */
  /* own_faces (nullable): int32 [nv][8], the faces incident to each cloth vertex
     (-1 padded; a vertex on more than eight faces has a row of -1).  The narrow
     phase skips a vertex's own faces (XPBD.py:258); given this table, the walks
     of the cloth tree by the cloth's vertices count them in a list's length (so
     a list overflows exactly when the reference's does, and keeps the entries it
     keeps) but store only the other candidates, in the same order.  Positions and
     velocities are the same with or without it; cand_n of that set then counts
     the stored entries.  (No reference counterpart: the reference's lists hold
     the incident faces too.) */
  const int32_t* own_faces;
  /* cache_every_static: substeps between rebuilds of the first query set's cache
     (cloth vertices against the static tree, which never moves: an entry stays
     valid while its query stays within the margin); 0 = cache_every */
  int32_t cache_every_static;
} slx_frame_desc;

size_t slx_frame_ws_bytes(const slx_frame_desc* d);
int slx_frame(const slx_frame_desc* d, slx_stream_t stream);

/* A solver's own side streams and events (default priority, created on the
 * current device), so that frames of different solvers never share them. */
int slx_frame_ctx_create(slx_frame_ctx** out);
int slx_frame_ctx_destroy(slx_frame_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
