// One frame of the cloth step, orchestrated on the host side of the library:
// the sequence of Solver.forward (framework/physics/XPBD.py:443-469).
//
//  if collision: rebuild the dynamic tree from x (and the static tree when asked)
//  per substep:  [triangle tree: refit from x, on a side stream, joined before
//                 the first query of the dynamic tree; the static-tree queries
//                 run on another side stream beside the dynamic-tree queries]
//                spring pass with compute_y fused in
//                per iteration: [spring pass] broad phase (3 query sets, and
//                               the edge sets when on), collision pass (the
//                               last one records the contacting pairs when the
//                               velocity pass is on)
//                compute_velocity + update_x, or with enable_velocity_update
//                compute_velocity, velocity pass, update_x (XPBD.py:465-469)
//  without collision (and without the velocity pass) the substeps run as two
//  fused kernels per substep (slx_substeps_plain).
//
// Every size the kernels need is either static or device-resident, so the
// frame enqueues work without any host synchronisation and can be captured in
// a HIP graph.
#include <mutex>

#include "slx_common.h"

extern "C" {
int slx_substeps_plain(int nv, float* x4, float* v4, float* x4_alt, float* v4_alt, float* y4a, float* y4b,
                       const float* fixed, const int32_t* slice_off, const int32_t* slice_width,
                       const int32_t* entries, int n_substeps, int n_iters, float dt, float gx, float gy, float gz,
                       float k_s, float k_b, float damping, slx_stream_t stream);
}

namespace {

using namespace slx;

bool batched(const slx_tree& t) { return t.mode == 1 && t.n_inst > 1; }

}  // namespace

// Two extra streams (fork/join with events, also under graph capture; default
// priority): s for the per-substep refit, which depends only on x and so runs
// beside the spring pass; q for the static-tree queries, which run beside the
// dynamic-tree queries.  Each solver owns a set (slx_frame_ctx); frames whose
// descriptor carries none share one set per device.
struct slx_frame_ctx {
  int device = -1;
  hipStream_t s = nullptr, q = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, qfork = nullptr, qjoin = nullptr, sub = nullptr, tvjoin = nullptr;
};

namespace {

using SideStream = slx_frame_ctx;

void destroy_side(SideStream& ss) {
  for (hipEvent_t* e : {&ss.fork, &ss.join, &ss.qfork, &ss.qjoin, &ss.sub, &ss.tvjoin})
    if (*e) {
      (void)hipEventDestroy(*e);
      *e = nullptr;
    }
  for (hipStream_t* q : {&ss.s, &ss.q})
    if (*q) {
      (void)hipStreamDestroy(*q);
      *q = nullptr;
    }
}

int create_side(SideStream& ss, int dev) {
  ss.device = dev;
  if (hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ss.q, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ss.join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ss.qfork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ss.qjoin, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ss.sub, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ss.tvjoin, hipEventDisableTiming) != hipSuccess) {
    destroy_side(ss);
    return fail(SLX_EHIP, "slx_frame: side stream creation failed");
  }
  return SLX_OK;
}

SideStream g_side[64];
std::mutex g_side_mu;

// the descriptor's context (on the current device), else the device's shared set
int side_stream(slx_frame_ctx* ctx, SideStream** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return fail(SLX_EHIP, "slx_frame: no device");
  if (ctx) {
    SLX_REQUIRE(ctx->device == dev, "slx_frame: the context belongs to device %d, the current device is %d",
                ctx->device, dev);
    *out = ctx;
    return SLX_OK;
  }
  std::lock_guard<std::mutex> lock(g_side_mu);
  SideStream& ss = g_side[dev];
  if (!ss.s) {
    int rc = create_side(ss, dev);
    if (rc) return rc;
  }
  *out = &ss;
  return SLX_OK;
}

size_t tree_ws(const slx_tree& t) {
  if (t.nf <= 0) return 0;
  if (batched(t)) return slx_seg_build_ws_bytes(t.nf, t.n_inst);
  return t.mode == 0 ? slx_cell_build_ws_bytes(t.nf) : slx_tri_build_ws_bytes(t.nf);
}

int build_tree(const slx_tree& t, int nv, const float* x4, const int32_t* faces, float pad, int32_t* status,
               void* ws, size_t ws_bytes, slx_stream_t s) {
  if (t.nf <= 0) return SLX_OK;
  if (batched(t))
    return slx_seg_build(nv, t.nf, t.n_inst, faces, x4, pad, t.vert_base, t.face_base, t.face_inst, t.fmin4, t.fmax4,
                         t.box6, t.codes, t.sorted_faces, t.sorted_inst, t.child_a, t.child_b, t.parent, t.range_l,
                         t.range_r, t.roots, t.sched, t.nlo4, t.nhi4, t.quad4, ws, ws_bytes, s);
  int rc = slx_face_aabbs(t.nf, faces, x4, pad, t.fmin4, t.fmax4, s);
  if (rc) return rc;
  rc = slx_mesh_aabb(nv, x4, pad, t.box6, ws, ws_bytes, s);
  if (rc) return rc;
  if (t.mode == 0)
    return slx_cell_build(t.nf, t.fmin4, t.fmax4, t.box6, t.child_a, t.child_b, t.parent, t.sched, t.face_cell,
                          t.prefix, t.sorted_faces, t.num, t.leaf_l, t.leaf_r, t.nlo4, t.nhi4, status, ws, ws_bytes, s);
  return slx_tri_build(t.nf, t.fmin4, t.fmax4, t.box6, t.codes, t.sorted_faces, t.child_a, t.child_b, t.parent,
                       t.range_l, t.range_r, t.sched, t.nlo4, t.nhi4, t.quad4, ws, ws_bytes, s);
}

// a cache struct with its redo list enables the two-tier walk; with leaves too,
// the cache proper
bool tiered(const slx_tree& t, const slx_cand_cache& c) {
  return t.mode == 1 && t.nf > 0 && c.redo != nullptr;
}
bool cacheable(const slx_tree& t, const slx_cand_cache& c) { return tiered(t, c) && c.cap > 0 && c.leaves; }

// queries of a mesh whose vertices belong to instances q_inst (batched trees only)
// q_order: processing order of the query vertices (nullable)
// cache (nullable): the set's cache struct; use_cache: filled when *built is
// false, then filtered; otherwise only its redo list serves the two-tier walk.
// disp: the tree's displacement since the cache was built (nullable = 0)
// own (nullable): the queries' own faces, left out of the stored lists
// redo_name: timer label of the second-tier walk
// (the frame keeps every redo counter at zero between passes: prezeroed walks)
int query(const char* name, const slx_tree& t, int nq, const float* q4, const int32_t* q_inst,
          const int32_t* q_order, float pad, int cap, int32_t* cand, int32_t* cand_n, int32_t* status,
          slx_stream_t s, const slx_cand_cache* cache = nullptr, bool use_cache = false, bool* built = nullptr,
          const uint32_t* disp = nullptr, const int32_t* own = nullptr, const char* redo_name = nullptr) {
  int nl = t.mode == 0 ? kCells : t.nf;
  if (t.nf <= 0) nl = 0;
  const bool seg = batched(t);
  if (cache && tiered(t, *cache)) {
    int kind = 0;
    if (use_cache && cacheable(t, *cache)) {
      kind = *built ? 2 : 1;
      *built = true;
    }
    return traverse_cached(name, kind, nq, q4, pad, nl, t.nlo4, t.nhi4, t.quad4, cache, disp, cap, cand, cand_n,
                           status, (hipStream_t)s, q_order, nullptr, true, nullptr, seg ? t.roots : nullptr,
                           seg ? q_inst : nullptr, own, redo_name);
  }
  return traverse_launch(name, t.mode, nq, q4, pad, nl, t.num, t.sorted_faces, t.nlo4, t.nhi4,
                         t.mode == 1 ? t.quad4 : nullptr, t.fmin4, t.fmax4, t.box6, cap, cand, cand_n, status,
                         (hipStream_t)s, seg ? t.roots : nullptr, seg ? q_inst : nullptr, q_order, nullptr, nullptr,
                         t.mode == 1 ? own : nullptr);
}

}  // namespace

extern "C" size_t slx_frame_ws_bytes(const slx_frame_desc* d) {
  if (!d) return 0;
  if (!d->collision) return 2 * sizeof(float4) * (size_t)(d->nv > 0 ? d->nv : 0);
  size_t w = slx_mesh_aabb_ws_bytes();
  if (d->collision) {
    size_t a = tree_ws(d->dyn), b = tree_ws(d->st);
    size_t c = slx_collision_ws_bytes(d->nv, d->nvs, d->rec_cap);
    if (a > w) w = a;
    if (b > w) w = b;
    if (c > w) w = c;
  }
  return w;
}

extern "C" int slx_frame_ctx_create(slx_frame_ctx** out) {
  SLX_REQUIRE(out, "slx_frame_ctx_create: null pointer");
  *out = nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(SLX_EHIP, "slx_frame_ctx_create: no device");
  slx_frame_ctx* c = new slx_frame_ctx();
  int rc = create_side(*c, dev);
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return SLX_OK;
}

extern "C" int slx_frame_ctx_destroy(slx_frame_ctx* ctx) {
  if (!ctx) return SLX_OK;
  destroy_side(*ctx);
  delete ctx;
  return SLX_OK;
}

static int frame_impl(const slx_frame_desc* d, slx_stream_t stream);

extern "C" int slx_frame(const slx_frame_desc* d, slx_stream_t stream) {
  SLX_REQUIRE(d, "slx_frame: null descriptor");
  SLX_REQUIRE(d->spring_block == 0 || d->spring_block == 64 || d->spring_block == 128 || d->spring_block == 256,
              "slx_frame: spring_block must be 0, 64, 128 or 256");
  SLX_REQUIRE(d->spring_vpt == 0 || d->spring_vpt == 1 || d->spring_vpt == 2 || d->spring_vpt == 4,
              "slx_frame: spring_vpt must be 0, 1, 2 or 4");
  SLX_REQUIRE(d->walk_stack == 0 || d->walk_stack == 16 || d->walk_stack == 24 || d->walk_stack == 32 ||
                  d->walk_stack == 64,
              "slx_frame: walk_stack must be 0, 16, 24, 32 or 64");
  SLX_REQUIRE(d->active_grid >= 0 && d->active_grid <= 65535, "slx_frame: active_grid outside 0..65535");
  SLX_REQUIRE(d->refit_sub >= 0 && d->refit_sub <= 2048 && d->refit_top >= 0 && d->refit_top <= 2048,
              "slx_frame: refit_sub / refit_top outside 0..2048");
  SLX_REQUIRE(d->walk_wave >= -1 && d->walk_wave <= 1, "slx_frame: walk_wave must be -1, 0 or 1");
  Knobs k = knobs();
  if (d->spring_block) k.spring_block = d->spring_block;
  if (d->spring_vpt) k.spring_vpt = d->spring_vpt;
  if (d->walk_stack) k.walk_stack = d->walk_stack;
  if (d->active_grid) k.active_grid = d->active_grid;
  if (d->refit_sub) k.refit_sub = d->refit_sub;
  if (d->refit_top) k.refit_top = d->refit_top;
  if (d->walk_wave) k.walk_wave = d->walk_wave > 0 ? 1 : 0;
  KnobScope scope(k);
  return frame_impl(d, stream);
}

static int frame_impl(const slx_frame_desc* d, slx_stream_t stream) {
  SLX_REQUIRE(d->nv >= 0 && d->nvs >= 0 && d->n_substeps >= 0 && d->n_iters >= 1, "slx_frame: bad sizes");
  if (d->nv == 0 || d->n_substeps == 0) return SLX_OK;
  SLX_REQUIRE(d->x4 && d->v4 && d->y4 && d->y4_alt && d->fixed && d->m_inv && d->adj_slice_off &&
                  d->adj_slice_width && d->adj_entries,
              "slx_frame: null state pointer");
  hipStream_t st = (hipStream_t)stream;
  SLX_REQUIRE(d->method == 0 || d->method == 1, "slx_frame: method must be 0 (Jacobi) or 1 (Gauss-Seidel)");
  SLX_REQUIRE(d->method == 0 || d->n_colors == 0 || (d->gs_edges && d->gs_color_off),
              "slx_frame: Gauss-Seidel colouring missing");
  const bool vel = d->velocity != 0;
  const bool sew = d->n_sewn > 0;
  SLX_REQUIRE(!sew || (d->sew_verts && d->sew_off && d->sew_list && d->sew_pairs && d->sew_ws),
              "slx_frame: sewing needs its pair lists and workspace");
  if (!d->collision && d->method == 0 && !vel && !sew) {
    // x4_alt / v4_alt: the caller's y4_alt doubles as one ping-pong buffer is not
    // possible (four buffers are needed); use the workspace for the alternates.
    size_t need = 2 * sizeof(float4) * (size_t)d->nv;
    SLX_REQUIRE(d->ws && d->ws_bytes >= need, "slx_frame: workspace too small for the collision-free path");
    float* xa = (float*)d->ws;
    float* va = xa + 4 * (size_t)d->nv;
    return slx_substeps_plain(d->nv, d->x4, d->v4, xa, va, d->y4, d->y4_alt, d->fixed, d->adj_slice_off,
                              d->adj_slice_width, d->adj_entries, d->n_substeps, d->n_iters, d->dt, d->gx, d->gy,
                              d->gz, d->k_s, d->k_b, d->damping, stream);
  }
  const bool coll = d->collision != 0;
  SLX_REQUIRE(!coll || (d->faces && d->cand && d->cand_n && d->ws), "slx_frame: null collision pointer");
  SLX_REQUIRE(!coll || d->dyn.mode == 0 || d->dyn.mode == 1, "slx_frame: bad tree mode");
  SLX_REQUIRE(!coll || d->nvs == 0 || (d->xs4 && d->faces_st), "slx_frame: static mesh missing");
  SLX_REQUIRE(!coll || !batched(d->dyn) || (d->dyn.vert_inst && d->dyn.roots && (d->nvs == 0 || (batched(d->st) &&
                                                                                               d->st.vert_inst))),
              "slx_frame: batched trees need instance ids on both meshes");
  SLX_REQUIRE(d->ws_bytes >= slx_frame_ws_bytes(d), "slx_frame: workspace too small");
  SLX_REQUIRE(!vel || !coll || (d->pair_cnt && d->pair_rec && d->pair_cap >= 0),
              "slx_frame: the velocity pass needs pair records");
  const bool ee = coll && d->ee.ne > 0;
  if (ee) {
    SLX_REQUIRE(!batched(d->dyn), "slx_frame: edge-edge constraints need a single (not batched) dynamic mesh");
    SLX_REQUIRE(d->ee.c_dy && d->ee.n_dy && (d->nvs == 0 || (d->ee.c_st && d->ee.n_st)),
                "slx_frame: edge-edge constraints need their candidate buffers");
    SLX_REQUIRE(!vel || (d->ee.pair_cnt && d->ee.pair_rec), "slx_frame: the velocity pass needs edge pair records");
  }
  const int nv = d->nv, nvs = coll ? d->nvs : 0, cap = d->cap;
  const float pad = d->padding;
  int rc;
  // the edge sets of a pass: static set only with a static mesh; pair records
  // only on the recording pass
  slx_edge_sets ee_rec = d->ee, ee_norec = d->ee;
  if (ee) {
    if (nvs == 0) ee_rec.c_st = ee_norec.c_st = ee_rec.n_st = ee_norec.n_st = nullptr;
    ee_norec.pair_cnt = nullptr;
    ee_norec.pair_rec = nullptr;
    if (!vel) ee_rec = ee_norec;
  }
  // the edge gate: triangle trees (their walks take a device list), static maxima once
  const bool gate_st = ee && d->ee.gate && d->ee.live && nvs > 0 && d->st.mode == 1 && !batched(d->st) &&
                       d->st.nf > 0 && d->ee.edges_st;
  const bool gate_dy = ee && d->ee.gate && d->ee.live && d->dyn.mode == 1 && d->dyn.nf > 0;
  int32_t* live_st = gate_st ? d->ee.live : nullptr;
  int32_t* live_dy_list = gate_dy ? d->ee.live + 1 + (size_t)d->ee.ne : nullptr;
  if (gate_st) {
    rc = ee_gate_static(&d->ee, d->ee.nes, (const float4*)d->xs4, st);
    if (rc) return rc;
  }
  // edge boxes of the current y against a tree (XPBD.py:217-232); triangle trees
  // take the two-tier walk when the caller gave redo lists (set 0 static, 1 dynamic),
  // over the gate's list when there is one
  auto ee_query = [&](const slx_tree& t, int set, const char* name, const float* y, const int32_t* c,
                      const int32_t* n, hipStream_t qs, const int32_t* live) {
    int nl = t.nf <= 0 ? 0 : (t.mode == 0 ? kCells : t.nf);
    if (t.mode == 1 && d->ee.redo && nl > 0) {
      slx_cand_cache walk{};
      walk.redo = d->ee.redo + (size_t)set * (1 + (size_t)d->ee.ne);
      // the coherent order pays for its scattered list writes on the deep dynamic tree only
      const int32_t* ord = live ? live + 1 : (set == 1 ? d->ee.order : nullptr);
      return traverse_cached(name, 0, d->ee.ne, y, pad, nl, t.nlo4, t.nhi4, t.quad4, &walk, nullptr, d->ee.cap,
                             (int32_t*)c, (int32_t*)n, d->status, qs, ord, d->ee.edges_dy, true, live);
    }
    if (live && nl > 0)
      return traverse_launch(name, 1, d->ee.ne, y, pad, nl, t.num, t.sorted_faces, t.nlo4, t.nhi4, t.quad4,
                             t.fmin4, t.fmax4, t.box6, d->ee.cap, (int32_t*)c, (int32_t*)n, d->status, qs, nullptr,
                             nullptr, live + 1, live, d->ee.edges_dy);
    return traverse_launch(name, t.mode, d->ee.ne, y, pad, nl, t.num, t.sorted_faces, t.nlo4,
                           t.nhi4, t.mode == 1 ? t.quad4 : nullptr, t.fmin4, t.fmax4, t.box6, d->ee.cap,
                           (int32_t*)c, (int32_t*)n, d->status, qs, nullptr, nullptr, nullptr, nullptr,
                           d->ee.edges_dy);
  };
  // per-frame rebuild from x (XPBD.py:448-458)
  if (coll) {
    rc = build_tree(d->dyn, nv, d->x4, d->faces, pad, d->status, d->ws, d->ws_bytes, stream);
    if (rc) return rc;
  }
  // candidate caches: rebuilt at each set's first query of the frame, against the
  // dynamic tree's vertices as they are now (x_snap) and its boxes as built
  // a cache pays off only when it is filtered at least once before its rebuild:
  // the cloth-vertex sets are queried n_iters times per substep, the
  // obstacle-vertex set once
  const bool reuse_v = d->cache_every != 1 || d->n_iters > 1, reuse_t = d->cache_every != 1;
  // the first set's cache (cloth vertices against the static tree) has an interval
  // of its own: that tree never moves (D = 0), so a query's entry stays valid
  // while the query itself stays within the margin
  const int every_st = d->cache_every_static != 0 ? d->cache_every_static : d->cache_every;
  const bool reuse_s = d->n_iters > 1 || (every_st != 1 && d->n_substeps > 1);
  const bool dyn_cached = coll && ((reuse_v && cacheable(d->dyn, d->cache[1])) ||
                                   (reuse_t && cacheable(d->dyn, d->cache[2])));
  SLX_REQUIRE(!dyn_cached || (d->dyn.x_snap && d->dyn.disp), "slx_frame: cached queries of the dynamic tree need "
              "its x_snap and disp");
  // caches rebuilt at every substep: x does not change between a rebuild and the
  // filters that use it (the same substep), so the tree's displacement since is
  // zero -- zeroed once below, with no snapshot and no measurement per substep
  const bool every_substep = d->cache_every == 1;
  bool built[3] = {false, false, false};
  if (nvs > 0 && d->rebuild_static) {
    rc = build_tree(d->st, nvs, d->xs4, d->faces_st, pad, d->status, d->ws, d->ws_bytes, stream);
    if (rc) return rc;
  }
  int32_t *c0 = d->cand, *c1 = d->cand + (size_t)cap * nv, *c2 = d->cand + 2 * (size_t)cap * nv;
  int32_t *n0 = d->cand_n, *n1 = d->cand_n + nv, *n2 = d->cand_n + 2 * (size_t)nv;
  int32_t* pn = d->pair_num;
  float4* ycur = (float4*)d->y4;
  float4* yalt = (float4*)d->y4_alt;
  bool tv_fresh = false;   // set-2 candidates depend only on the static x and the dynamic tree
  SideStream* side = nullptr;
  const bool refits = coll && d->dyn.mode == 1 && d->refit_every_substep && d->dyn.nf > 0 && d->n_substeps > 1;
  // every walk of a single dynamic tree reads its 4-wide lines (and leaves), unless
  // a query set is dense enough next to it for the binary wavefront walk (the
  // obstacle vertices with at least 8 per dynamic face): then the refit keeps the
  // binary node boxes too
  const bool lines_only = d->dyn.quad4 && !batched(d->dyn) && (long long)d->dyn.nf * 8 > (long long)nvs;
  // The buckets' counters and the walks' redo counters are zeroed once here (the
  // tree builds above reuse the workspace); from then on every apply kernel leaves
  // them at zero for the next pass, so no pass needs a zeroing launch of its own.
  ZeroList after{};
  if (coll) {
    rc = side_stream(d->ctx, &side);
    if (rc) return rc;
    auto add = [&](int32_t* p) {
      if (p) after.p[after.n++] = p;
    };
    for (int k = 0; k < 3; ++k) add(d->cache[k].redo);
    if (ee && d->ee.redo) {
      add(d->ee.redo);
      add(d->ee.redo + 1 + (size_t)d->ee.ne);
    }
    if (dyn_cached && every_substep) add((int32_t*)d->dyn.disp);
    rc = frame_zero(nv, d->rec_cap, d->ws, after, st);
    if (rc) return rc;
  }
  // Substeps that only filter the dynamic caches need the tree's leaf boxes from x,
  // not its internal boxes: the filters test cached leaves, and the few queries
  // they send to a full walk walk the internal boxes of the last full refit (at the
  // caches' rebuild, from x_snap) with their box grown by the displacement bound D
  // since, testing leaves by their current boxes -- the same leaves in the same
  // order.  Those substeps refit the leaves only.
  const bool stale_ok = refits && dyn_cached && !every_substep && !ee && !batched(d->dyn) && lines_only &&
                        (long long)d->dyn.nf * 8 > (long long)nv;
  bool tree_stale = false;
  for (int s = 0; s < d->n_substeps; ++s) {
    bool refit_pending = false, snapped = false, disp_pending = false, v_done = false;
    const bool rebuild_now = coll && (s == 0 || (d->cache_every > 0 && s % d->cache_every == 0));
    if (refits && s > 0) {
      // fork: the refit reads x as the previous substep left it (ahead of the
      // snapshot below, which it does not need)
      if (hipEventRecord(side->fork, st) != hipSuccess || hipStreamWaitEvent(side->s, side->fork, 0) != hipSuccess)
        return fail(SLX_EHIP, "slx_frame: fork failed");
      tree_stale = stale_ok && !rebuild_now;
      if (tree_stale) {
        rc = tri_leaf_boxes_impl(d->dyn.nf, d->faces, d->x4, pad, d->dyn.sorted_faces, d->dyn.nlo4, d->dyn.nhi4,
                                 side->s);
        if (!rc && hipEventRecord(side->sub, side->s) != hipSuccess) return fail(SLX_EHIP, "slx_frame: event failed");
      } else {
        rc = tri_refit_impl(d->dyn.nf, d->faces, d->x4, pad, d->dyn.sorted_faces, d->dyn.child_a, d->dyn.child_b,
                            d->dyn.parent, d->dyn.sched, d->dyn.nlo4, d->dyn.nhi4, d->dyn.quad4, side->s, side->sub,
                            lines_only);
      }
      if (rc) return rc;
      disp_pending = dyn_cached && !every_substep;   // measured on the main stream after the first spring pass
      if (hipEventRecord(side->join, side->s) != hipSuccess) return fail(SLX_EHIP, "slx_frame: join failed");
      refit_pending = true;
      tv_fresh = false;
    }
    if (coll && (s == 0 || (every_st > 0 && s % every_st == 0))) built[0] = false;
    if (rebuild_now) {
      // caches rebuilt in this substep, against x as it is now (the refit above
      // builds the dynamic tree's boxes from this x)
      built[1] = built[2] = false;
      if (dyn_cached && !every_substep) {
        rc = slx_tree_snapshot(nv, d->x4, d->dyn.x_snap, d->dyn.disp, stream);   // (zeroes disp)
        if (rc) return rc;
        snapped = true;
      }
    }
    auto join_refit = [&]() -> int {
      if (!refit_pending) return SLX_OK;
      refit_pending = false;
      return hipStreamWaitEvent(st, side->join, 0) == hipSuccess ? SLX_OK : fail(SLX_EHIP, "slx_frame: join failed");
    };
    if (d->method == 1) {
      rc = slx_compute_y(nv, d->x4, d->v4, (float*)ycur, d->dt, d->gx, d->gy, d->gz, stream);
      if (rc) return rc;
    }
    for (int it = 0; it < d->n_iters; ++it) {
      if (d->method == 0) {
        // Jacobi spring pass (compute_y fused into the first)
        rc = spring_launch(it == 0, false, nv, (const float4*)d->x4, (const float4*)d->v4, d->fixed, ycur, yalt,
                           nullptr, nullptr, d->adj_slice_off, d->adj_slice_width, d->adj_entries, d->k_s, d->k_b,
                           d->dt, d->gx, d->gy, d->gz, d->damping, st);
        if (rc) return rc;
        float4* t = ycur; ycur = yalt; yalt = t;
      } else {
        rc = gs_pass(nv, (const float4*)d->x4, d->fixed, ycur, d->adj_slice_off, d->adj_slice_width, d->adj_entries,
                     (const int4*)d->gs_edges, d->gs_color_off, d->n_colors, d->k_s, d->k_b, st);
        if (rc) return rc;
      }
      if (!coll) {
        if (sew) {
          rc = slx_sewing_pass(d->n_sewn, d->sew_verts, d->sew_off, d->sew_list, d->sew_pairs, (float*)ycur, d->fixed,
                               d->k_sew, d->sew_ws, stream);
          if (rc) return rc;
        }
        continue;
      }
      // broad phase (XPBD.py:186-215); the static-tree queries, and the narrow
      // phase of the first query set (which needs only their lists and only
      // moves its own vertex), on their own stream beside the dynamic tree's
      // queries
      const bool rec = vel && it == d->n_iters - 1;   // the velocity pass reads the last pass's pairs
      // the edge gate of this pass: the dynamic set only when the dynamic tree's
      // boxes come from the current x (refit this substep, or built this frame)
      const int32_t* live_dy = (gate_dy && (refits || s == 0)) ? live_dy_list : nullptr;
      if (ee && d->ee.gate && d->ee.live) {    // (also clears the gate's flags when neither set is gated)
        rc = ee_gate(&d->ee, nv, (const float4*)d->x4, ycur, pad, it == 0, live_st, (int32_t*)live_dy, st);
        if (rc) return rc;
      }
      if (disp_pending) {                    // how far the refit boxes' vertices moved since the caches
        rc = tree_displacement(nv, d->x4, d->dyn.x_snap, d->dyn.disp, !snapped, st);
        if (rc) return rc;
        disp_pending = false;
      }
      bool static_pending = false, tv_side = false;
      // the static tree's queries and the first set's contacts on the query stream
      // part: 0 everything, 1 the obstacle vertices' query alone, 2 everything else
      auto static_block = [&](int part) -> int {
        int rc2 = SLX_OK;
        if (nvs > 0) {
          if (hipEventRecord(side->qfork, st) != hipSuccess || hipStreamWaitEvent(side->q, side->qfork, 0) != hipSuccess)
            return fail(SLX_EHIP, "slx_frame: fork failed");
          // and after the refit's subtree kernel: beside it their workgroups hold
          // the CUs whose whole LDS each refit workgroup needs (refit 333 -> 203
          // us, frame -4.7%); its top part is one workgroup
          if (refit_pending && hipStreamWaitEvent(side->q, side->sub, 0) != hipSuccess)
            return fail(SLX_EHIP, "slx_frame: fork failed");
          if (part != 1) {
            rc2 = query("traverse_vt_st", d->st, nv, (const float*)ycur, d->dyn.vert_inst, d->dyn.vert_order, pad, cap,
                        c0, n0, d->status, (slx_stream_t)side->q, &d->cache[0], reuse_s, &built[0], nullptr, nullptr,
                        "walk_redo_vt_st");
            if (rc2) return rc2;
          }
          if (part != 2 && !tv_fresh) {        // obstacle vertices vs the dynamic tree, beside its cache build
            if (refit_pending && hipStreamWaitEvent(side->q, side->join, 0) != hipSuccess)
              return fail(SLX_EHIP, "slx_frame: fork failed");
            Knobs kst = knobs();
            kst.stale_disp = tree_stale ? d->dyn.disp : nullptr;
            KnobScope ks_stale(kst);
            rc2 = query("traverse_tv_st", d->dyn, nvs, d->xs4, d->st.vert_inst, d->st.vert_order, pad, cap, c2, n2,
                       d->status, (slx_stream_t)side->q, &d->cache[2], reuse_t, &built[2], d->dyn.disp, nullptr,
                       "walk_redo_tv_st");
            if (rc2) return rc2;
            if (hipEventRecord(side->tvjoin, side->q) != hipSuccess) return fail(SLX_EHIP, "slx_frame: join failed");
            tv_fresh = tv_side = true;
          }
          if (part == 1) return rc2;
          rc2 = collision_pass_impl(nv, nvs, d->faces, d->faces_st, (float*)ycur, d->xs4, d->fixed, d->m_inv, d->dhat,
                                   d->k_col, cap, c0, n0, c1, n1, c2, n2, pn, rec ? d->pair_cap : 0,
                                   rec ? d->pair_cnt : nullptr, rec ? d->pair_rec : nullptr, nullptr, d->rec_cap,
                                   d->status, d->ws, d->ws_bytes, side->q, true, after, 1, nullptr);
          if (rc2) return rc2;
          if (ee) {
            rc2 = ee_query(d->st, 0, "traverse_ee_st", (const float*)ycur, d->ee.c_st, d->ee.n_st, side->q, live_st);
            if (rc2) return rc2;
          }
          if (hipEventRecord(side->qjoin, side->q) != hipSuccess) return fail(SLX_EHIP, "slx_frame: join failed");
          static_pending = true;
        }
        return rc2;
      };
      // Later passes of a substep fork it after the cloth's own cache filter: both
      // filters at once each took twice as long, on the critical path (1M drape
      // 19.27-19.33 -> 19.09 ms per frame).
      // A pass that only filters the cloth's caches (a leaf-only refit) splits it:
      // the obstacle vertices' query first (the main stream's contacts wait for it),
      // the rest after the cloth's filter.
      const bool static_after = it > 0;
      const bool static_split = !static_after && tree_stale;
      if (!static_after) {
        rc = static_block(static_split ? 1 : 0);
        if (rc) return rc;
      }
      rc = join_refit();                     // the dynamic tree is needed from here on
      if (rc) return rc;
      {
        Knobs kst = knobs();
        kst.stale_disp = tree_stale ? d->dyn.disp : nullptr;
        KnobScope ks_stale(kst);
        rc = query("traverse_vt_dy", d->dyn, nv, (const float*)ycur, d->dyn.vert_inst, d->dyn.vert_order, pad, cap,
                   c1, n1, d->status, stream, &d->cache[1], reuse_v, &built[1], d->dyn.disp, d->own_faces,
                   "walk_redo_vt_dy");
      }
      if (rc) return rc;
      if (static_after || static_split) {
        rc = static_block(static_split ? 2 : 0);
        if (rc) return rc;
      }
      if (tv_side && hipStreamWaitEvent(st, side->tvjoin, 0) != hipSuccess)
        return fail(SLX_EHIP, "slx_frame: join failed");
      if (ee) {
        rc = ee_query(d->dyn, 1, "traverse_ee_dy", (const float*)ycur, d->ee.c_dy, d->ee.n_dy, st, live_dy);
        if (rc) return rc;
      }
      // narrow phase + projection + update_dx (XPBD.py:236-281, 365-369): the
      // other two sets here, then (after the static stream's join) the edge sets
      // and the apply; the substep's last apply also computes the velocity when
      // nothing moves y after it
      const bool fuse_v = vel && !sew && it == d->n_iters - 1;
      const VelFuse vfuse{(const float4*)d->x4, (float4*)d->v4, d->damping, d->dt};
      v_done = v_done || fuse_v;
      rc = collision_pass_impl(nv, nvs, d->faces, d->faces_st, (float*)ycur, d->xs4, d->fixed, d->m_inv, d->dhat,
                               d->k_col, cap, nvs > 0 ? c0 : nullptr, n0, c1, n1, nvs > 0 ? c2 : nullptr, n2, pn,
                               rec ? d->pair_cap : 0, rec ? d->pair_cnt : nullptr, rec ? d->pair_rec : nullptr,
                               ee ? (rec ? &ee_rec : &ee_norec) : nullptr, d->rec_cap, d->status, d->ws,
                               d->ws_bytes, st, true, after, static_pending ? 2 : 0,
                               static_pending ? side->qjoin : nullptr, fuse_v ? &vfuse : nullptr);
      if (rc) return rc;
      if (sew) {                             // sewing constraints (XPBD.py:371-372, there commented out)
        rc = slx_sewing_pass(d->n_sewn, d->sew_verts, d->sew_off, d->sew_list, d->sew_pairs, (float*)ycur, d->fixed,
                             d->k_sew, d->sew_ws, stream);
        if (rc) return rc;
      }
    }
    rc = join_refit();                       // (n_iters == 0: nothing queried the tree)
    if (rc) return rc;
    if (vel) {
      if (!v_done) {
        rc = slx_compute_velocity(nv, d->x4, d->v4, (const float*)ycur, d->fixed, d->damping, d->dt, stream);
        if (rc) return rc;
      }
      rc = velocity_pass_impl(nv, nvs, d->faces, d->faces_st, d->v4, d->vs4, d->fixed, d->m_inv, d->m_inv_st,
                              d->mu, d->pair_cap, coll ? d->pair_cnt : nullptr, d->pair_rec,
                              ee ? &ee_rec : nullptr, d->x4, d->dt, d->rec_cap > 0 ? d->rec_cap : 1, d->status,
                              d->ws, d->ws_bytes, st, coll);
    } else {
      rc = slx_update_vx(nv, d->x4, d->v4, (const float*)ycur, d->fixed, d->damping, d->dt, stream);
    }
    if (rc) return rc;
  }
  if (ycur != (float4*)d->y4) {
    if (hipMemcpyAsync(d->y4, ycur, sizeof(float4) * (size_t)nv, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return fail(SLX_EHIP, "slx_frame: y copy failed");
  }
  return SLX_OK;
}
