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
// dynamic triangles) first counts its contact records, an exclusive scan turns
// the counts into record offsets in that serial order, the records
// (target vertex, delta) are written, stably sorted by target vertex, and each
// vertex sums its records in order.  The result equals a serial scatter in the
// reference's loop order, bit for bit, with no float atomics.
#include "slx_common.h"

namespace slx {

size_t radix_sort_ws_bytes(int n_max);
int radix_sort_pairs(int n_max, const int* n_dev, uint32_t* keys, int32_t* vals, int key_bits,
                     void* ws, size_t ws_bytes, hipStream_t st);
size_t scan_ws_bytes(int n_max);
int exclusive_scan(int n_max, const int* n_dev, const int* in, int* out, int* total_dev, void* ws,
                   size_t ws_bytes, hipStream_t st);

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

// squared distance of the region's primitive pair and its gradient; g[0..3] =
// (p, t0, t1, t2), zero for points outside the pair
__device__ float dist_grad(int dtype, float3 p, float3 t0, float3 t1, float3 t2, float3* g) {
  const float3 z = make_float3(0.f, 0.f, 0.f);
  g[0] = g[1] = g[2] = g[3] = z;
  if (dtype <= 2) {                      // point-point: d = |p-a|^2, dd/dp = 2(p-a)
    float3 a = dtype == 0 ? t0 : (dtype == 1 ? t1 : t2);
    float3 d = sub3(p, a);
    float3 gp = make_float3(2.0f * d.x, 2.0f * d.y, 2.0f * d.z);
    g[0] = gp;
    g[1 + dtype] = make_float3(-gp.x, -gp.y, -gp.z);
    return dot3(d, d);
  }
  if (dtype <= 5) {                      // point-edge (a, b)
    int ia = dtype == 3 ? 1 : (dtype == 4 ? 2 : 1);
    int ib = dtype == 3 ? 2 : 3;
    float3 a = ia == 1 ? t0 : t1, b = ib == 2 ? t1 : t2;
    float3 u = sub3(a, p), w = sub3(b, p), c = cross3(u, w), e = sub3(b, a);
    float L = dot3(e, e);
    float d = dot3(c, c) / L;
    float3 wc = cross3(w, c), cu = cross3(c, u);
    float twoL = 2.0f / L;
    float q = twoL * d;
    float3 ga = make_float3(twoL * wc.x + q * e.x, twoL * wc.y + q * e.y, twoL * wc.z + q * e.z);
    float3 gb = make_float3(twoL * cu.x - q * e.x, twoL * cu.y - q * e.y, twoL * cu.z - q * e.z);
    g[ia] = ga;
    g[ib] = gb;
    g[0] = make_float3(-(ga.x + gb.x), -(ga.y + gb.y), -(ga.z + gb.z));
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
  g[0] = gp;
  g[2] = gb;
  g[3] = gc;
  g[1] = make_float3(-((gp.x + gb.x) + gc.x), -((gp.y + gb.y) + gc.y), -((gp.z + gb.z) + gc.z));
  return d;
}

__device__ __constant__ unsigned char c_slots[7] = {0x3, 0x5, 0x9, 0x7, 0xD, 0xB, 0xF};

struct Contact {
  int vid[4];
  float3 delta[4];
  unsigned mask;      // slots that receive a record
  bool record;        // pair recorded for the velocity pass
};

// One candidate of query set `set` (0 vt_st, 1 vt_dy, 2 tv_st); fills the
// records in slot order (collision_constraints_x.py:5-310).
__device__ void contact(int set, int qv, int f, const int32_t* fdy, const int32_t* fst, const float4* y4,
                        const float4* xs4, const float* fixed, const float* m_inv, float dhat, float k,
                        Contact& c) {
  c.mask = 0;
  c.record = false;
  float3 P[4];
  if (set == 0) {
    c.vid[0] = qv; c.vid[1] = fst[3 * f]; c.vid[2] = fst[3 * f + 1]; c.vid[3] = fst[3 * f + 2];
    P[0] = xyz(y4[qv]); P[1] = xyz(xs4[c.vid[1]]); P[2] = xyz(xs4[c.vid[2]]); P[3] = xyz(xs4[c.vid[3]]);
  } else if (set == 1) {
    c.vid[0] = qv; c.vid[1] = fdy[3 * f]; c.vid[2] = fdy[3 * f + 1]; c.vid[3] = fdy[3 * f + 2];
    P[0] = xyz(y4[qv]); P[1] = xyz(y4[c.vid[1]]); P[2] = xyz(y4[c.vid[2]]); P[3] = xyz(y4[c.vid[3]]);
  } else {
    c.vid[0] = qv; c.vid[1] = fdy[3 * f]; c.vid[2] = fdy[3 * f + 1]; c.vid[3] = fdy[3 * f + 2];
    P[0] = xyz(xs4[qv]); P[1] = xyz(y4[c.vid[1]]); P[2] = xyz(y4[c.vid[2]]); P[3] = xyz(y4[c.vid[3]]);
  }
  const int dtype = d_type_pt(P[0], P[1], P[2], P[3]);
  float3 g[4];
  const float d = dist_grad(dtype, P[0], P[1], P[2], P[3], g);
  const unsigned m = c_slots[dtype];
  if (set == 0) {                                        // __vt_st
    if (d < dhat) {
      float w = m_inv[qv];
      float schur = w * dot3(g[0], g[0]);
      float ld = (k * (dhat - d)) / ((k * schur) + 1.0f);
      c.delta[0] = mul3(w * ld, g[0]);
      c.mask = 1u;
      c.record = true;
    }
    return;
  }
  c.record = true;                                       // __tv_st / __vt_dy record every pair
  if (!(d < dhat)) return;
  const int s0 = (set == 2) ? 1 : 0;                     // static point does not move
  float schur = 0.f;
  bool first = true;
  for (int s = s0; s < 4; ++s) {
    if (!(m & (1u << s))) continue;
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
  for (int s = s0; s < 4; ++s) {
    if (!(m & (1u << s))) continue;
    int v = c.vid[s];
    float a = (set == 1 && dtype == 0) ? m_inv[v] * ld : (fixed[v] * m_inv[v]) * ld;
    c.delta[s] = mul3(a, g[s]);
    c.mask |= 1u << s;
  }
}

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

// Pass 1: number of contact records per query (and recorded pair counts).
// Records of the first set (dynamic vertex vs static triangle) only ever move
// the query vertex itself and come first in its serial order, so that thread
// sums them in candidate order right here (self_acc = (sum, count)); only the
// other two sets go through the emit / sort path.
__global__ void k_contact_count(QuerySets Q, const int32_t* __restrict__ fdy, const int32_t* __restrict__ fst,
                                const float4* __restrict__ y4, const float4* __restrict__ xs4,
                                const float* __restrict__ fixed, const float* __restrict__ m_inv, float dhat,
                                float k, int32_t* __restrict__ cnt, int32_t* __restrict__ pair_num,
                                float4* __restrict__ self_acc) {
  const int nqt = 2 * Q.nv + Q.nvs;
  int qg = blockIdx.x * blockDim.x + threadIdx.x;
  if (qg >= nqt) return;
  int set, qv, ncand, nq;
  const int32_t* cand;
  int recs = 0, pairs = 0;
  float3 dx = make_float3(0.f, 0.f, 0.f);
  if (query_of(Q, qg, set, qv, cand, ncand, nq)) {
    for (int j = 0; j < ncand; ++j) {
      int f = cand[(size_t)j * nq + qv];
      if (set == 1 && (fdy[3 * f] == qv || fdy[3 * f + 1] == qv || fdy[3 * f + 2] == qv)) continue;
      Contact c;
      contact(set, qv, f, fdy, fst, y4, xs4, fixed, m_inv, dhat, k, c);
      if (set == 0 && c.mask) dx = make_float3(dx.x + c.delta[0].x, dx.y + c.delta[0].y, dx.z + c.delta[0].z);
      recs += __popc(c.mask);
      pairs += c.record ? 1 : 0;
    }
  }
  if (qg < Q.nv) {
    self_acc[qg] = make_float4(dx.x, dx.y, dx.z, (float)recs);
    recs = 0;
  }
  cnt[qg] = recs;
  if (pair_num) pair_num[qg] = pairs < Q.cap ? pairs : Q.cap;
}

// Pass 2: write the records at their serial-order offsets.
__global__ void k_contact_emit(QuerySets Q, const int32_t* __restrict__ fdy, const int32_t* __restrict__ fst,
                               const float4* __restrict__ y4, const float4* __restrict__ xs4,
                               const float* __restrict__ fixed, const float* __restrict__ m_inv, float dhat,
                               float k, const int32_t* __restrict__ off, int rec_cap, uint32_t* __restrict__ rkey,
                               int32_t* __restrict__ ridx, float4* __restrict__ rval, int32_t* status) {
  const int nqt = 2 * Q.nv + Q.nvs;
  int qg = Q.nv + blockIdx.x * blockDim.x + threadIdx.x;     // the first set is summed in pass 1
  if (qg >= nqt) return;
  int set, qv, ncand, nq;
  const int32_t* cand;
  if (!query_of(Q, qg, set, qv, cand, ncand, nq)) return;
  int o = off[qg];
  for (int j = 0; j < ncand; ++j) {
    int f = cand[(size_t)j * nq + qv];
    if (set == 1 && (fdy[3 * f] == qv || fdy[3 * f + 1] == qv || fdy[3 * f + 2] == qv)) continue;
    Contact c;
    contact(set, qv, f, fdy, fst, y4, xs4, fixed, m_inv, dhat, k, c);
    for (int s = 0; s < 4; ++s) {
      if (!(c.mask & (1u << s))) continue;
      if (o < rec_cap) {
        rkey[o] = (uint32_t)c.vid[s];
        ridx[o] = o;
        rval[o] = make_float4(c.delta[s].x, c.delta[s].y, c.delta[s].z, 0.f);
      } else if (status) {
        atomicOr(status, kStatusCandOverflow);
      }
      ++o;
    }
  }
}

__global__ void k_clamp_total(int32_t* total, int rec_cap) {
  if (threadIdx.x == 0 && *total > rec_cap) *total = rec_cap;
}

// segment bounds of each vertex in the sorted record keys
__global__ void k_seg_bounds(int rec_cap, const int32_t* __restrict__ total, const uint32_t* __restrict__ key,
                             int32_t* __restrict__ seg_lo, int32_t* __restrict__ seg_hi) {
  const int n = *total;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t k = key[i];
  if (i == 0 || key[i - 1] != k) seg_lo[k] = i;
  if (i == n - 1 || key[i + 1] != k) seg_hi[k] = i + 1;
}

// update_dx over the contact records: y += f * (dx / nc) (XPBD.py:383-387)
__global__ void k_contact_apply(int nv, const int32_t* __restrict__ seg_lo, const int32_t* __restrict__ seg_hi,
                                const int32_t* __restrict__ ridx, const float4* __restrict__ rval,
                                const float4* __restrict__ self_acc, const float* __restrict__ fixed,
                                float4* __restrict__ y4) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const float4 sa = self_acc[v];            // first-set records, already summed in order
  int lo = seg_lo[v], hi = seg_hi[v];
  if (hi < lo) hi = lo;
  const float nc = sa.w + (float)(hi - lo);
  if (!(nc > 0.f)) return;
  float3 dx = make_float3(sa.x, sa.y, sa.z);
  for (int i = lo; i < hi; ++i) {
    float4 r = rval[ridx[i]];
    dx = make_float3(dx.x + r.x, dx.y + r.y, dx.z + r.z);
  }
  float f = fixed[v];
  float4 y = y4[v];
  y4[v] = make_float4(y.x + f * (dx.x / nc), y.y + f * (dx.y / nc), y.z + f * (dx.z / nc), y.w);
}

__global__ void k_fill2(int n, int32_t* __restrict__ a, int32_t* __restrict__ b) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { a[i] = 0; b[i] = 0; }
}

static int vbits(int nv) {
  int b = 1;
  while ((1ll << b) < (long long)nv) ++b;
  return b;
}

struct CollisionWs {
  int32_t *cnt, *off, *total, *seg_lo, *seg_hi, *ridx;
  uint32_t* rkey;
  float4* rval;
  float4* self_acc;
  void* scan_ws;
  size_t scan_bytes;
  void* sort_ws;
  size_t sort_bytes;
};

static size_t layout(int nv, int nvs, int rec_cap, char* base, CollisionWs* w) {
  size_t nq = (size_t)(2 * nv + nvs), o = 0;
  auto take = [&](size_t bytes) {
    size_t at = o;
    o += (bytes + 255) & ~(size_t)255;
    return base ? base + at : nullptr;
  };
  size_t rc = (size_t)(rec_cap > 0 ? rec_cap : 1);
  CollisionWs t;
  t.cnt = (int32_t*)take(4 * nq);
  t.off = (int32_t*)take(4 * nq);
  t.total = (int32_t*)take(256);
  t.seg_lo = (int32_t*)take(4 * (size_t)nv);
  t.seg_hi = (int32_t*)take(4 * (size_t)nv);
  t.rkey = (uint32_t*)take(4 * rc);
  t.ridx = (int32_t*)take(4 * rc);
  t.rval = (float4*)take(16 * rc);
  t.self_acc = (float4*)take(16 * (size_t)nv);
  t.scan_bytes = scan_ws_bytes((int)nq);
  t.scan_ws = take(t.scan_bytes);
  t.sort_bytes = radix_sort_ws_bytes((int)rc);
  t.sort_ws = take(t.sort_bytes);
  if (w) *w = t;
  return o;
}

}  // namespace slx

using namespace slx;

extern "C" size_t slx_collision_ws_bytes(int nv, int nvs, int rec_cap) {
  return layout(nv, nvs, rec_cap, nullptr, nullptr);
}

extern "C" int slx_collision_pass(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st, float* y4,
                                  const float* xs4, const float* fixed, const float* m_inv, float dhat, float k_col,
                                  int cap, const int32_t* c_vt_st, const int32_t* n_vt_st, const int32_t* c_vt_dy,
                                  const int32_t* n_vt_dy, const int32_t* c_tv_st, const int32_t* n_tv_st,
                                  int32_t* pair_num, int rec_cap, int32_t* status, void* ws, size_t ws_bytes,
                                  slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0 && nvs >= 0 && cap >= 0 && rec_cap >= 0, "slx_collision_pass: bad sizes");
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(faces_dy && y4 && fixed && m_inv && ws, "slx_collision_pass: null pointer");
  SLX_REQUIRE(!(c_vt_st || c_tv_st) || (faces_st && xs4), "slx_collision_pass: static mesh missing");
  SLX_REQUIRE((!c_vt_st || n_vt_st) && (!c_vt_dy || n_vt_dy) && (!c_tv_st || n_tv_st),
              "slx_collision_pass: candidate counts missing");
  if (ws_bytes < slx_collision_ws_bytes(nv, nvs, rec_cap))
    return fail(SLX_EWS, "slx_collision_pass: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  CollisionWs w;
  layout(nv, nvs, rec_cap, (char*)ws, &w);
  QuerySets Q{nv, c_tv_st ? nvs : 0, cap, c_vt_st, n_vt_st, c_vt_dy, n_vt_dy, c_tv_st, n_tv_st};
  const int nqt = 2 * nv + Q.nvs;
  SLX_LAUNCH("contact_count", k_contact_count, dim3(blocks_for(nqt, 256)), dim3(256), 0, st, Q, faces_dy, faces_st,
             (const float4*)y4, (const float4*)xs4, fixed, m_inv, dhat, k_col, w.cnt, pair_num, w.self_acc);
  int rc = exclusive_scan(nqt, nullptr, w.cnt, w.off, w.total, w.scan_ws, w.scan_bytes, st);
  if (rc) return rc;
  SLX_LAUNCH("contact_emit", k_contact_emit, dim3(blocks_for(nqt - nv, 256)), dim3(256), 0, st, Q, faces_dy, faces_st,
             (const float4*)y4, (const float4*)xs4, fixed, m_inv, dhat, k_col, w.off, rec_cap, w.rkey, w.ridx,
             w.rval, status);
  SLX_LAUNCH("clamp_total", k_clamp_total, dim3(1), dim3(64), 0, st, w.total, rec_cap);
  if (rec_cap > 0) {
    rc = radix_sort_pairs(rec_cap, w.total, w.rkey, w.ridx, vbits(nv), w.sort_ws, w.sort_bytes, st);
    if (rc) return rc;
  }
  SLX_LAUNCH("fill", k_fill2, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, w.seg_lo, w.seg_hi);
  if (rec_cap > 0)
    SLX_LAUNCH("seg_bounds", k_seg_bounds, dim3(blocks_for(rec_cap, 256)), dim3(256), 0, st, rec_cap, w.total,
               w.rkey, w.seg_lo, w.seg_hi);
  SLX_LAUNCH("contact_apply", k_contact_apply, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, w.seg_lo, w.seg_hi,
             w.ridx, w.rval, w.self_acc, fixed, (float4*)y4);
  return check_launch("slx_collision_pass");
}
