// Narrow phase and collision projection on gfx950 (vertex-triangle
// non-penetration constraints), with deterministic accumulation.
//
// Reference: framework/physics/XPBD.py:236-264 (solve_collision_constraints_x),
// framework/physics/collision_constraints_x.py:5-310 (__vt_st, __tv_st,
// __vt_dy), framework/collision/distance.py:4-116 (region classification and
// squared distances; gradients derived here), XPBD.py:383-387 (update_dx).
//
// The reference scatters each contact's deltas with float atomics.  Here every
// query (the three sets in the reference's loop order: dynamic vertices vs
// static triangles, dynamic vertices vs dynamic triangles, static vertices vs
// dynamic triangles) computes its contacts; each delta becomes a record in its
// target vertex's bucket tagged with its position in that serial order, and each
// vertex adds its records in that order.  The result equals a serial scatter in
// the reference's loop order, bit for bit, with no float atomics.
#include "slx_common.h"

namespace slx {

// ---- distance.py restated (same op order as the CPU oracle) ----

__device__ __forceinline__ void edge_param(float3 p, float3 a, float3 b, float3 n, float& s, float& t) {
  float3 e = sub3(b, a), en = cross3(e, n), r = sub3(p, a);
  float m00 = dot3(e, e), m01 = dot3(e, en), m10 = dot3(en, e), m11 = dot3(en, en);
  float det = m00 * m11 - m01 * m10;
  float id = 1.0f / det;
  float i00 = id * m11, i01 = id * (-m01), i10 = id * (-m10), i11 = id * m00;
  float b0 = dot3(e, r), b1 = dot3(en, r);
  s = i00 * b0 + i01 * b1;
  t = i10 * b0 + i11 * b1;
}

// d_type_PT, distance.py:4-47
__device__ int d_type_pt(float3 p, float3 t0, float3 t1, float3 t2) {
  float3 n = cross3(sub3(t1, t0), sub3(t2, t0));
  float s0, u0, s1, u1, s2, u2;
  edge_param(p, t0, t1, n, s0, u0);
  if (s0 > 0.f && s0 < 1.f && u0 >= 0.f) return 3;
  edge_param(p, t1, t2, n, s1, u1);
  if (s1 > 0.f && s1 < 1.f && u1 >= 0.f) return 4;
  edge_param(p, t2, t0, n, s2, u2);
  if (s2 > 0.f && s2 < 1.f && u2 >= 0.f) return 5;
  if (s0 <= 0.f && s2 >= 1.f) return 0;
  if (s1 <= 0.f && s0 >= 1.f) return 1;
  if (s2 <= 0.f && s1 >= 1.f) return 2;
  return 6;
}

// squared distance of the region's primitive pair and its gradient g0..g3 for
// (p, t0, t1, t2), zero for points outside the pair
__device__ __forceinline__ float dist_grad(int dtype, float3 p, float3 t0, float3 t1, float3 t2, float3& g0,
                                           float3& g1, float3& g2, float3& g3) {
  const float3 z = make_float3(0.f, 0.f, 0.f);
  g0 = g1 = g2 = g3 = z;
  if (dtype <= 2) {                      // point-point: d = |p-a|^2, dd/dp = 2(p-a)
    float3 a = dtype == 0 ? t0 : (dtype == 1 ? t1 : t2);
    float3 d = sub3(p, a);
    float3 gp = make_float3(2.0f * d.x, 2.0f * d.y, 2.0f * d.z);
    float3 ga = make_float3(-gp.x, -gp.y, -gp.z);
    g0 = gp;
    if (dtype == 0) g1 = ga;
    else if (dtype == 1) g2 = ga;
    else g3 = ga;
    return dot3(d, d);
  }
  if (dtype <= 5) {                      // point-edge (a, b): 3 = (t0,t1), 4 = (t1,t2), 5 = (t0,t2)
    float3 a = dtype == 4 ? t1 : t0, b = dtype == 3 ? t1 : t2;
    float3 u = sub3(a, p), w = sub3(b, p), c = cross3(u, w), e = sub3(b, a);
    float L = dot3(e, e);
    float d = dot3(c, c) / L;
    float3 wc = cross3(w, c), cu = cross3(c, u);
    float twoL = 2.0f / L;
    float q = twoL * d;
    float3 ga = make_float3(twoL * wc.x + q * e.x, twoL * wc.y + q * e.y, twoL * wc.z + q * e.z);
    float3 gb = make_float3(twoL * cu.x - q * e.x, twoL * cu.y - q * e.y, twoL * cu.z - q * e.z);
    if (dtype == 3) { g1 = ga; g2 = gb; }
    else if (dtype == 4) { g2 = ga; g3 = gb; }
    else { g1 = ga; g3 = gb; }
    g0 = make_float3(-(ga.x + gb.x), -(ga.y + gb.y), -(ga.z + gb.z));
    return d;
  }
  // point-plane: d = ((p-a).n)^2/|n|^2
  float3 e1 = sub3(t1, t0), e2 = sub3(t2, t0), n = cross3(e1, e2), r = sub3(p, t0);
  float s = dot3(r, n), N = dot3(n, n);
  float d = (s * s) / N;
  float q = s / N;
  float q2 = 2.0f * q, qq2 = (2.0f * q) * q;
  float3 e2r = cross3(e2, r), e2n = cross3(e2, n), re1 = cross3(r, e1), ne1 = cross3(n, e1);
  float3 gp = make_float3(q2 * n.x, q2 * n.y, q2 * n.z);
  float3 gb = make_float3(q2 * e2r.x - qq2 * e2n.x, q2 * e2r.y - qq2 * e2n.y, q2 * e2r.z - qq2 * e2n.z);
  float3 gc = make_float3(q2 * re1.x - qq2 * ne1.x, q2 * re1.y - qq2 * ne1.y, q2 * re1.z - qq2 * ne1.z);
  g0 = gp;
  g2 = gb;
  g3 = gc;
  g1 = make_float3(-((gp.x + gb.x) + gc.x), -((gp.y + gb.y) + gc.y), -((gp.z + gb.z) + gc.z));
  return d;
}

__device__ __constant__ unsigned char c_slots[7] = {0x3, 0x5, 0x9, 0x7, 0xD, 0xB, 0xF};

// One candidate's deltas for slots (p, t0, t1, t2); every loop over the four slots
// is unrolled, so the arrays stay in registers.
struct Contact {
  int vid[4];
  float3 delta[4];
  unsigned mask;      // slots that receive a delta
  bool record;        // pair recorded for the velocity pass
  bool hit;           // d < dhat (a recorded pair that can act on velocities)
  int dtype;
  float3 g[4];        // gradients as recorded (zero outside the pair's points)
  float schur;        // as recorded
};

// One candidate of query set `set` (0 vt_st, 1 vt_dy, 2 tv_st), slots in order
// (collision_constraints_x.py:5-310).
__device__ __forceinline__ void contact(int set, int qv, int3 fv, const float4* y4, const float4* xs4,
                                        const float* fixed, const float* m_inv, float dhat, float k, Contact& c) {
  c.mask = 0;
  c.record = false;
  c.hit = false;
  c.schur = 0.f;
  c.vid[0] = qv;
  c.vid[1] = fv.x;
  c.vid[2] = fv.y;
  c.vid[3] = fv.z;
  const float4* tp = set == 0 ? xs4 : y4;
  const float3 p = xyz(set == 2 ? xs4[qv] : y4[qv]);
  const float3 t0 = xyz(tp[c.vid[1]]), t1 = xyz(tp[c.vid[2]]), t2 = xyz(tp[c.vid[3]]);
  const int dtype = d_type_pt(p, t0, t1, t2);
  float3 g[4];
  const float d = dist_grad(dtype, p, t0, t1, t2, g[0], g[1], g[2], g[3]);
  const unsigned m = c_slots[dtype];
  c.dtype = dtype;
#pragma unroll
  for (int s = 0; s < 4; ++s) c.g[s] = g[s];
  if (set == 0) {                                        // __vt_st
    if (d < dhat) {
      float w = m_inv[qv];
      float schur = w * dot3(g[0], g[0]);
      float ld = (k * (dhat - d)) / ((k * schur) + 1.0f);
      c.delta[0] = mul3(w * ld, g[0]);
      c.mask = 1u;
      c.record = true;
      c.hit = true;
      c.schur = schur;
    }
    return;
  }
  c.record = true;                                       // __tv_st / __vt_dy record every pair
  if (!(d < dhat)) return;
  c.hit = true;
  const int s0 = (set == 2) ? 1 : 0;                     // static point does not move
  float schur = 0.f;
  bool first = true;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s < s0 || !(m & (1u << s))) continue;
    int v = c.vid[s];
    float term = (fixed[v] * m_inv[v]) * dot3(g[s], g[s]);
    schur = first ? term : schur + term;
    first = false;
  }
  float ld;
  if (set == 2) {
    ld = (k * (dhat - d)) / ((k * schur) + 1.0f);
  } else {
    schur = schur + 1e-4f;
    ld = (dhat - d) / schur;
  }
  c.schur = schur;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    if (s < s0 || !(m & (1u << s))) continue;
    int v = c.vid[s];
    float a = (set == 1 && dtype == 0) ? m_inv[v] * ld : (fixed[v] * m_inv[v]) * ld;
    c.delta[s] = mul3(a, g[s]);
    c.mask |= 1u << s;
  }
}

// Candidates are loaded kCandGroup at a time.  Most vertices of the draped 1M
// cloth have one to three stored candidates, and the kernel's occupancy is set by
// its registers: one at a time, with the register budget of five waves per SIMD
// (a few bytes of spills), gave 19.30-19.32 ms per frame against 19.46-19.52 ms
// for groups of four at four waves (six waves spilled too much: 20.9 ms).
constexpr int kCandGroup = 1;

struct QuerySets {
  int nv, nvs, cap;
  const int32_t *c0, *n0, *c1, *n1, *c2, *n2;  // column-major candidate lists [cap][nq]
};

__device__ __forceinline__ bool query_of(const QuerySets& Q, int qg, int& set, int& qv, const int32_t*& cand,
                                         int& ncand, int& nq) {
  if (qg < Q.nv) { set = 0; qv = qg; cand = Q.c0; ncand = Q.c0 ? Q.n0[qv] : 0; nq = Q.nv; }
  else if (qg < 2 * Q.nv) { set = 1; qv = qg - Q.nv; cand = Q.c1; ncand = Q.c1 ? Q.n1[qv] : 0; nq = Q.nv; }
  else { set = 2; qv = qg - 2 * Q.nv; cand = Q.c2; ncand = Q.c2 ? Q.n2[qv] : 0; nq = Q.nvs; }
  if (ncand > Q.cap) ncand = Q.cap;
  return cand != nullptr;
}

// Every record (target vertex, delta) carries its position in the reference's
// serial loop order, o = ((query - nv) * cap + candidate) * 4 + slot over the
// second and third query sets (the first set only ever moves its own query
// vertex and comes first, so its thread sums those deltas itself).  A record
// claims a slot of its target vertex's bucket with an integer atomic; the vertex
// then adds its records in increasing o.  The sum is therefore the serial
// scatter's, bit for bit, whatever order the slots were claimed in.
//
// REC: the query also stores its contacting pair records for the velocity pass
// (see slx_collision_pass for the layout).
template <bool REC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
k_contacts(QuerySets Q, const int32_t* __restrict__ fdy, const int32_t* __restrict__ fst,
           const float4* __restrict__ y4, const float4* __restrict__ xs4, const float* __restrict__ fixed,
           const float* __restrict__ m_inv, float dhat, float k, int bucket, int32_t* __restrict__ bcnt,
           float4* __restrict__ brec, float4* __restrict__ self_acc, int32_t* __restrict__ pair_num,
           int pcap, int32_t* __restrict__ pair_cnt, float4* __restrict__ pair_rec, int32_t* status, int q_begin,
           int q_end) {
  const int nqt = 2 * Q.nv + Q.nvs;
  // (in dispatch order, not xcd_block(): grouping its blocks by XCD measured
  // slower on the 1M-vertex drape, 119 -> 126 us for the first set)
  const int qg = q_begin + (int)(blockIdx.x * blockDim.x + threadIdx.x);   // queries [q_begin, q_end)
  if (qg >= q_end) return;
  int set, qv, ncand, nq;
  const int32_t* cand;
  int pairs = 0, own = 0, stored = 0;
  bool over = false, pover = false;
  float3 dx = make_float3(0.f, 0.f, 0.f);
  if (query_of(Q, qg, set, qv, cand, ncand, nq)) {
    const int32_t* F = set == 0 ? fst : fdy;
    // candidates in groups of kCandGroup: all ids and corner ids of a group are
    // loaded before it is processed, in candidate order
    for (int j0 = 0; j0 < ncand; j0 += kCandGroup) {
      int3 fv[kCandGroup];
      int fid[kCandGroup];
#pragma unroll
      for (int g = 0; g < kCandGroup; ++g) {
        const int f = j0 + g < ncand ? cand[(size_t)(j0 + g) * nq + qv] : -1;
        fid[g] = f;
        fv[g] = f >= 0 ? make_int3(F[3 * f], F[3 * f + 1], F[3 * f + 2]) : make_int3(-1, -1, -1);
      }
#pragma unroll
      for (int g = 0; g < kCandGroup; ++g) {
        if (j0 + g >= ncand) break;
        if (set == 1 && (fv[g].x == qv || fv[g].y == qv || fv[g].z == qv)) continue;
        Contact c;
        contact(set, qv, fv[g], y4, xs4, fixed, m_inv, dhat, k, c);
        if (REC && c.record && c.hit && pairs < Q.cap) {
          if (stored < pcap) {
            const size_t col = (size_t)stored * 4u * (size_t)nqt + (size_t)qg;
            pair_rec[col] = make_float4(c.g[0].x, c.g[0].y, c.g[0].z, c.schur);
            pair_rec[col + (size_t)nqt] = make_float4(c.g[1].x, c.g[1].y, c.g[1].z, __int_as_float(fid[g]));
            pair_rec[col + 2 * (size_t)nqt] = make_float4(c.g[2].x, c.g[2].y, c.g[2].z, __int_as_float(c.dtype));
            pair_rec[col + 3 * (size_t)nqt] = make_float4(c.g[3].x, c.g[3].y, c.g[3].z, 0.f);
            ++stored;
          } else {
            pover = true;
          }
        }
        pairs += c.record ? 1 : 0;
        if (set == 0) {
          if (c.mask) {
            dx = make_float3(dx.x + c.delta[0].x, dx.y + c.delta[0].y, dx.z + c.delta[0].z);
            ++own;
          }
          continue;
        }
        const uint32_t o0 = ((uint32_t)(qg - Q.nv) * (uint32_t)Q.cap + (uint32_t)(j0 + g)) * 4u;
#pragma unroll
        for (int sl = 0; sl < 4; ++sl) {
          if (!(c.mask & (1u << sl))) continue;
          const int t = c.vid[sl];
          const int at = atomicAdd(&bcnt[t], 1);
          if (at < bucket)
            brec[(size_t)t * bucket + at] = make_float4(c.delta[sl].x, c.delta[sl].y, c.delta[sl].z,
                                                        __uint_as_float(o0 + (uint32_t)sl));
          else
            over = true;
        }
      }
    }
  }
  if (qg < Q.nv) self_acc[qg] = make_float4(dx.x, dx.y, dx.z, (float)own);
  if (pair_num) pair_num[qg] = pairs < Q.cap ? pairs : Q.cap;
  if (REC) pair_cnt[qg] = stored;
  if (over && status) atomicOr(status, kStatusRecOverflow);
  if (REC && pover && status) atomicOr(status, kStatusPairOverflow);
}

// update_dx over the contacts: y += f * (dx / nc) (XPBD.py:383-387); the bucket
// is summed in increasing serial order.  The bucket's counter is left at zero for
// the next pass, and so are the counters in `after`.
// VEL: also compute_velocity from the final y (the substep's last pass).
template <bool VEL>
__global__ void __launch_bounds__(256)
k_contact_apply(int nv, int bucket, int32_t* __restrict__ bcnt, const float4* __restrict__ brec,
                const float4* __restrict__ self_acc, const float* __restrict__ fixed, float4* __restrict__ y4,
                ZeroList after, VelFuse vf) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && (int)threadIdx.x < after.n) *after.p[threadIdx.x] = 0;
  if (v >= nv) return;
  const float4 sa = self_acc[v];            // first-set deltas, already summed in order
  int n = bcnt[v];
  if (n != 0) bcnt[v] = 0;
  if (n > bucket) n = bucket;
  const float nc = sa.w + (float)n;
  if (!VEL && !(nc > 0.f)) return;
  const float f = fixed[v];
  float4 y = y4[v];
  if (nc > 0.f) {
    const float3 dx = bucket_sum(make_float3(sa.x, sa.y, sa.z), brec + (size_t)v * bucket, n);
    y = make_float4(y.x + f * (dx.x / nc), y.y + f * (dx.y / nc), y.z + f * (dx.z / nc), y.w);
    y4[v] = y;
  }
  if (VEL) {                                // compute_velocity on the final y
    const float4 x = vf.x4[v];
    const float af = (1.0f - vf.damping) * f;
    vf.v4[v] = make_float4((af * (y.x - x.x)) / vf.dt, (af * (y.y - x.y)) / vf.dt, (af * (y.z - x.z)) / vf.dt, f);
  }
}

__global__ void k_zero_i32(int n, int32_t* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_zero_i32_list(int n, int32_t* __restrict__ a, ZeroList zl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && (int)threadIdx.x < zl.n) *zl.p[threadIdx.x] = 0;
  if (i < n) a[i] = 0;
}

}  // namespace slx

using namespace slx;

extern "C" size_t slx_collision_ws_bytes(int nv, int nvs, int rec_cap) {
  (void)nvs;
  return contact_ws_layout(nv, rec_cap, nullptr, nullptr);
}

namespace slx {
int frame_zero(int nv, int rec_cap, void* ws, const ZeroList& zl, hipStream_t st) {
  ContactWs w;
  contact_ws_layout(nv, rec_cap, (char*)ws, &w);
  SLX_LAUNCH("fill", k_zero_i32_list, dim3(blocks_for(nv > 0 ? nv : 1, 256)), dim3(256), 0, st, nv, w.bcnt, zl);
  return check_launch("frame_zero");
}

int collision_pass_impl(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st, float* y4,
                        const float* xs4, const float* fixed, const float* m_inv, float dhat, float k_col, int cap,
                        const int32_t* c_vt_st, const int32_t* n_vt_st, const int32_t* c_vt_dy,
                        const int32_t* n_vt_dy, const int32_t* c_tv_st, const int32_t* n_tv_st, int32_t* pair_num,
                        int pair_cap, int32_t* pair_cnt, float* pair_rec, const slx_edge_sets* ee, int rec_cap,
                        int32_t* status, void* ws, size_t ws_bytes, hipStream_t st, bool bcnt_zeroed,
                        const ZeroList& after, int set0, hipEvent_t before_apply, const VelFuse* vel) {
  SLX_REQUIRE(nv >= 0 && nvs >= 0 && cap >= 0 && rec_cap >= 1, "slx_collision_pass: bad sizes");
  SLX_REQUIRE(set0 >= 0 && set0 <= 2, "slx_collision_pass: bad first-set mode");
  if (ee) {
    int rc = ee_check(ee, false);
    if (rc) return rc;
  }
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(faces_dy && y4 && fixed && m_inv && ws, "slx_collision_pass: null pointer");
  SLX_REQUIRE(!(c_vt_st || c_tv_st) || (faces_st && xs4), "slx_collision_pass: static mesh missing");
  SLX_REQUIRE((!c_vt_st || n_vt_st) && (!c_vt_dy || n_vt_dy) && (!c_tv_st || n_tv_st),
              "slx_collision_pass: candidate counts missing");
  SLX_REQUIRE((long long)(nv + nvs) * (long long)cap * 4ll < (1ll << 32),
              "slx_collision_pass: too many queries x candidates for 32-bit order keys");
  const bool rec = pair_rec != nullptr;
  SLX_REQUIRE(!rec || (pair_cnt && pair_cap >= 0), "slx_collision_pass: pair records need pair_cnt and pair_cap");
  if (ws_bytes < slx_collision_ws_bytes(nv, nvs, rec_cap))
    return fail(SLX_EWS, "slx_collision_pass: workspace too small");
  ContactWs w;
  contact_ws_layout(nv, rec_cap, (char*)ws, &w);
  QuerySets Q{nv, nvs, cap, c_vt_st, n_vt_st, c_vt_dy, n_vt_dy, c_tv_st, n_tv_st};
  const int nqt = 2 * nv + Q.nvs;
  // set0: 0 = all query sets here; 1 = only the first set (its deltas stay in
  // self_acc; nothing else is launched); 2 = the other two sets and the rest of
  // the pass (the first set launched before, on another stream, which
  // before_apply joins ahead of the edge sets and the apply)
  const int q_begin = set0 == 2 ? nv : 0, q_end = set0 == 1 ? nv : nqt;
  if (set0 != 1 && !bcnt_zeroed)
    SLX_LAUNCH("fill", k_zero_i32, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, w.bcnt);
  const char* name = set0 == 1 ? "contacts_vt_st" : "contacts";
  if (rec)
    SLX_LAUNCH(name, k_contacts<true>, dim3(blocks_for(q_end - q_begin, 256)), dim3(256), 0, st, Q, faces_dy,
               faces_st, (const float4*)y4, (const float4*)xs4, fixed, m_inv, dhat, k_col, rec_cap, w.bcnt, w.brec,
               w.self_acc, pair_num, pair_cap, pair_cnt, (float4*)pair_rec, status, q_begin, q_end);
  else
    SLX_LAUNCH(name, k_contacts<false>, dim3(blocks_for(q_end - q_begin, 256)), dim3(256), 0, st, Q, faces_dy,
               faces_st, (const float4*)y4, (const float4*)xs4, fixed, m_inv, dhat, k_col, rec_cap, w.bcnt, w.brec,
               w.self_acc, pair_num, 0, nullptr, nullptr, status, q_begin, q_end);
  if (set0 == 1) return check_launch("slx_collision_pass");
  if (before_apply && hipStreamWaitEvent(st, before_apply, 0) != hipSuccess)
    return fail(SLX_EHIP, "slx_collision_pass: join failed");
  if (ee) {
    int rc = ee_position_launch(nv, ee, (const float4*)y4, (const float4*)xs4, fixed, m_inv, dhat, k_col, w, rec_cap,
                                status, st);
    if (rc) return rc;
  }
  if (vel)
    SLX_LAUNCH("contact_apply", k_contact_apply<true>, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, rec_cap,
               w.bcnt, w.brec, w.self_acc, fixed, (float4*)y4, after, *vel);
  else
    SLX_LAUNCH("contact_apply", k_contact_apply<false>, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, rec_cap,
               w.bcnt, w.brec, w.self_acc, fixed, (float4*)y4, after, VelFuse{});
  return check_launch("slx_collision_pass");
}
}  // namespace slx

extern "C" int slx_collision_pass_ee(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st, float* y4,
                                     const float* xs4, const float* fixed, const float* m_inv, float dhat,
                                     float k_col, int cap, const int32_t* c_vt_st, const int32_t* n_vt_st,
                                     const int32_t* c_vt_dy, const int32_t* n_vt_dy, const int32_t* c_tv_st,
                                     const int32_t* n_tv_st, int32_t* pair_num, int pair_cap, int32_t* pair_cnt,
                                     float* pair_rec, const slx_edge_sets* ee, int rec_cap, int32_t* status,
                                     void* ws, size_t ws_bytes, slx_stream_t stream) {
  slx_edge_sets plain;                       // the gate is slx_frame's: a standalone pass tests every edge
  if (ee) {
    plain = *ee;
    plain.gate = nullptr;
    plain.live = nullptr;
  }
  return collision_pass_impl(nv, nvs, faces_dy, faces_st, y4, xs4, fixed, m_inv, dhat, k_col, cap, c_vt_st, n_vt_st,
                             c_vt_dy, n_vt_dy, c_tv_st, n_tv_st, pair_num, pair_cap, pair_cnt, pair_rec,
                             ee ? &plain : nullptr, rec_cap, status, ws, ws_bytes, (hipStream_t)stream, false,
                             ZeroList{}, 0, nullptr);
}

extern "C" int slx_collision_pass(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st, float* y4,
                                  const float* xs4, const float* fixed, const float* m_inv, float dhat, float k_col,
                                  int cap, const int32_t* c_vt_st, const int32_t* n_vt_st, const int32_t* c_vt_dy,
                                  const int32_t* n_vt_dy, const int32_t* c_tv_st, const int32_t* n_tv_st,
                                  int32_t* pair_num, int pair_cap, int32_t* pair_cnt, float* pair_rec,
                                  int rec_cap, int32_t* status, void* ws, size_t ws_bytes, slx_stream_t stream) {
  return slx_collision_pass_ee(nv, nvs, faces_dy, faces_st, y4, xs4, fixed, m_inv, dhat, k_col, cap, c_vt_st, n_vt_st,
                               c_vt_dy, n_vt_dy, c_tv_st, n_tv_st, pair_num, pair_cap, pair_cnt, pair_rec, nullptr,
                               rec_cap, status, ws, ws_bytes, stream);
}
