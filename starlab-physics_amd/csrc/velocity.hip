// Velocity-level contact and friction pass on gfx950, with deterministic
// accumulation.
//
// Reference: framework/physics/XPBD.py:374-381 (solve_constraints_v), 283-311
// (solve_collision_constraints_v), 397-404 (update_dv), 340-343
// (compute_velocity), 332-337 (update_x); framework/physics/
// collision_constraints_v.py:5-570 (__vt_st, __tv_st, __vt_dy).
//
// The position pass leaves, per query, the pairs that were in contact (their
// gradients g0..g3 and schur as recorded, slx_collision_pass).  One thread per
// query walks its pairs in record order; a pair with Cv = sum g_k . v_k < 0 gives
// each of its moving points a normal delta and then a friction delta.  As in the
// position pass, the first query set only moves its own vertex, so that thread
// sums its deltas; the other two sets put each delta in its target vertex's
// bucket with the key ((query - nv) * pair_cap + pair) * 8 + phase * 4 + slot,
// and the vertex adds its bucket in key order -- the reference's serial loop
// order -- before update_dv.  No float atomics.
#include "slx_common.h"

namespace slx {

__device__ __forceinline__ float norm3(float3 a) { return sqrtf(dot3(a, a)); }

// slots with deltas per contact type among (p, t0, t1, t2)
__device__ __constant__ unsigned char c_vslots[7] = {0x3, 0x5, 0x9, 0x7, 0xD, 0xB, 0xF};

// Normal (dn) and friction (dt) deltas of one pair, per slot in `mask`.
struct VelContact {
  int vid[4];
  float3 dn[4], dt[4];
  unsigned mask;
};

__device__ __forceinline__ float3 zero3() { return make_float3(0.f, 0.f, 0.f); }

// One recorded pair of query set `set` (0 vt_st, 1 vt_dy, 2 tv_st); returns
// false when Cv >= 0 (no deltas).  Op order as the serial restatement.
__device__ __forceinline__ bool vel_pair(int set, int qv, int3 fv, int dtype, const float3 (&g)[4], float schur,
                                         const float4* __restrict__ v4, const float4* __restrict__ vs4,
                                         const float* __restrict__ fixed, const float* __restrict__ m_inv,
                                         const float* __restrict__ m_inv_st, int nvs, float mu, VelContact& c) {
  c.vid[0] = qv;
  c.vid[1] = fv.x;
  c.vid[2] = fv.y;
  c.vid[3] = fv.z;
  float3 u[4];
  u[0] = set == 2 ? (vs4 ? xyz(vs4[qv]) : zero3()) : xyz(v4[qv]);
#pragma unroll
  for (int s = 1; s < 4; ++s) u[s] = set == 0 ? (vs4 ? xyz(vs4[c.vid[s]]) : zero3()) : xyz(v4[c.vid[s]]);
  // __tv_st takes g3 for t1 as well (collision_constraints_v.py:144)
  const float3 g2v = set == 2 ? g[3] : g[2];
  const float Cv = ((dot3(g[0], u[0]) + dot3(g[1], u[1])) + dot3(g2v, u[2])) + dot3(g[3], u[3]);
  if (!(Cv < 0.f)) return false;
  const float ld = -Cv / schur;
  const float mu_abs = mu * fabsf(Cv);

  if (set == 0) {                                         // __vt_st, :5-134
    const float m0 = m_inv[qv];
    float mn = m0;
    if (dtype == 3) mn = (m_inv_st && qv < nvs) ? m_inv_st[qv] : 0.f;
    c.dn[0] = mul3(mn * ld, g[0]);
    const float3 vt = add3(u[0], mul3(m0 * ld, g[0]));
    float3 p;
    if (dtype <= 2) {
      p = dtype == 0 ? u[1] : (dtype == 1 ? u[2] : u[3]);
    } else if (dtype <= 5) {
      const float3 ga = dtype == 4 ? g[2] : g[1], gb = dtype == 3 ? g[2] : g[3];
      const float3 qa = dtype == 4 ? u[2] : u[1], qb = dtype == 3 ? u[2] : u[3];
      const float a = norm3(ga), b = norm3(gb);
      const float ab = a + b;
      p = make_float3((a * qa.x + b * qb.x) / ab, (a * qa.y + b * qb.y) / ab, (a * qa.z + b * qb.z) / ab);
    } else {
      const float a = norm3(g[1]), b = norm3(g[2]), cc = norm3(g[3]);
      const float abc = (a + b) + cc;
      p = make_float3(((a * u[1].x + b * u[2].x) + cc * u[3].x) / abc,
                      ((a * u[1].y + b * u[2].y) + cc * u[3].y) / abc,
                      ((a * u[1].z + b * u[2].z) + cc * u[3].z) / abc);
    }
    const float3 gt = dtype == 5 ? sub3(vt, p) : sub3(p, vt);   // type 5: vTan0 - p (:109)
    const float dd = dot3(gt, gt);
    const float cT = 0.5f * dd;
    const float sT = m0 * dd + 1e-4f;
    const float lT = cT / sT;
    const float mul = mu_abs > cT ? 1.0f : mu;
    c.dt[0] = mul3(mul, mul3(m0 * lT, gt));
    c.mask = 1u;
    return true;
  }

  const unsigned m = set == 2 ? (c_vslots[dtype] & 0xEu) : c_vslots[dtype];
  c.mask = m;
  float3 vt[4], gt[4];
  float mt[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    vt[s] = zero3();
    gt[s] = zero3();
    mt[s] = 0.f;
    c.dn[s] = zero3();
    c.dt[s] = zero3();
    if (!(m & (1u << s))) continue;
    const int w = c.vid[s];
    const float fm = fixed[w] * m_inv[w];
    const float3 b = make_float3((fm * g[s].x) * ld, (fm * g[s].y) * ld, (fm * g[s].z) * ld);
    // normal delta: f w ld g (__tv_st) or f w g ld (__vt_dy)
    c.dn[s] = set == 2 ? mul3(fm * ld, g[s]) : b;
    vt[s] = add3(u[s], b);
    mt[s] = m_inv[w];
  }
  float cT = 0.f, sT = 1.f;
  if (set == 2) {                                         // __tv_st, :136-326
    if (dtype <= 2) {
      const int s = dtype + 1;
#pragma unroll
      for (int k = 1; k < 4; ++k)
        if (k == s) {
          gt[k] = sub3(u[0], vt[k]);
          const float dd = dot3(gt[k], gt[k]);
          cT = 0.5f * dd;
          sT = mt[k] * dd + 1e-4f;
        }
    } else if (dtype <= 5) {
      const int sa = dtype == 4 ? 2 : 1, sb = dtype == 3 ? 2 : 3;
      const float3 va = sa == 1 ? vt[1] : vt[2], vb = sb == 2 ? vt[2] : vt[3];
      const float a = norm3(sa == 1 ? g[1] : g[2]), b = norm3(sb == 2 ? g[2] : g[3]);
      const float ab = a + b;
      const float3 q = make_float3(u[0].x - (a * va.x + b * vb.x) / ab, u[0].y - (a * va.y + b * vb.y) / ab,
                                   u[0].z - (a * va.z + b * vb.z) / ab);
      const float3 ta = mul3(a / ab, q), tb = mul3(b / ab, q);
      const float ma = sa == 1 ? mt[1] : mt[2], mb = sb == 2 ? mt[2] : mt[3];
      cT = 0.5f * dot3(q, q);
      sT = (ma * dot3(ta, ta) + mb * dot3(tb, tb)) + 1e-4f;
#pragma unroll
      for (int k = 1; k < 4; ++k) {
        if (k == sa) gt[k] = ta;
        if (k == sb) gt[k] = tb;
      }
    } else {
      const float a = norm3(g[1]), b = norm3(g[2]), cc = norm3(g[3]);
      const float abc = (a + b) + cc;
      const float3 q = make_float3(u[0].x - ((a * vt[1].x + b * vt[2].x) + cc * vt[3].x) / abc,
                                   u[0].y - ((a * vt[1].y + b * vt[2].y) + cc * vt[3].y) / abc,
                                   u[0].z - ((a * vt[1].z + b * vt[2].z) + cc * vt[3].z) / abc);
      gt[1] = mul3(a / abc, q);
      gt[2] = mul3(b / abc, q);
      gt[3] = mul3(cc / abc, q);
      cT = 0.5f * dot3(q, q);
      sT = ((mt[1] * dot3(gt[1], gt[1]) + mt[2] * dot3(gt[2], gt[2])) + mt[3] * dot3(gt[3], gt[3])) + 1e-4f;
    }
  } else {                                                // __vt_dy, :329-570
    if (dtype <= 2) {
      const int s = dtype + 1;
      float3 vs_ = zero3();
      float ms = 0.f;
#pragma unroll
      for (int k = 1; k < 4; ++k)
        if (k == s) {
          vs_ = vt[k];
          ms = mt[k];
        }
      const float3 r = sub3(vt[0], vs_);
      const float3 g0t = make_float3(-r.x, -r.y, -r.z);
      const float d0 = dot3(g0t, g0t), d1 = dot3(r, r);
      cT = 0.5f * (d1 + d0);
      sT = (mt[0] * d0 + ms * d1) + 1e-4f;
      gt[0] = g0t;
#pragma unroll
      for (int k = 1; k < 4; ++k)
        if (k == s) gt[k] = r;
      if (dtype == 1) mt[2] = m_inv[c.vid[1]];          // t1's friction with t0's inverse mass (:388)
    } else if (dtype <= 5) {
      const int sa = dtype == 4 ? 2 : 1, sb = dtype == 3 ? 2 : 3;
      const float3 va = sa == 1 ? vt[1] : vt[2], vb = sb == 2 ? vt[2] : vt[3];
      const float a = norm3(sa == 1 ? g[1] : g[2]), b = norm3(sb == 2 ? g[2] : g[3]);
      const float ab = a + b;
      const float3 g0t = make_float3((a * va.x + b * vb.x) / ab - vt[0].x, (a * va.y + b * vb.y) / ab - vt[0].y,
                                     (a * va.z + b * vb.z) / ab - vt[0].z);
      const float3 ta = mul3(-(a / ab), g0t), tb = mul3(-(b / ab), g0t);
      const float ma = sa == 1 ? mt[1] : mt[2], mb = sb == 2 ? mt[2] : mt[3];
      const float d0 = dot3(g0t, g0t), da = dot3(ta, ta), db = dot3(tb, tb);
      cT = 0.5f * ((d0 + da) + db);
      sT = ((mt[0] * d0 + ma * da) + mb * db) + 1e-4f;
      gt[0] = g0t;
#pragma unroll
      for (int k = 1; k < 4; ++k) {
        if (k == sa) gt[k] = ta;
        if (k == sb) gt[k] = tb;
      }
    } else {
      const float a = norm3(g[1]), b = norm3(g[2]), cc = norm3(g[3]);
      const float abc = (a + b) + cc;
      const float3 g0t = make_float3(((a * vt[1].x + b * vt[2].x) + cc * vt[3].x) / abc - vt[0].x,
                                     ((a * vt[1].y + b * vt[2].y) + cc * vt[3].y) / abc - vt[0].y,
                                     ((a * vt[1].z + b * vt[2].z) + cc * vt[3].z) / abc - vt[0].z);
      gt[0] = g0t;
      gt[1] = mul3(-(a / abc), g0t);
      gt[2] = mul3(-(b / abc), g0t);
      gt[3] = mul3(-(cc / abc), g0t);
      const float d0 = dot3(gt[0], gt[0]), d1 = dot3(gt[1], gt[1]), d2 = dot3(gt[2], gt[2]),
                  d3 = dot3(gt[3], gt[3]);
      cT = 0.5f * (((d0 + d1) + d2) + d3);
      sT = (((mt[0] * d0 + mt[1] * d1) + mt[2] * d2) + mt[3] * d3) + 1e-4f;
    }
  }
  const float lT = cT / sT;
  const float mul = mu_abs > cT ? 1.0f : mu;
#pragma unroll
  for (int s = 0; s < 4; ++s)
    if (m & (1u << s)) c.dt[s] = mul3(mul, mul3(mt[s] * lT, gt[s]));
  return true;
}

__global__ void __launch_bounds__(256)
k_vel_contacts(int nv, int nvs, const int32_t* __restrict__ fdy, const int32_t* __restrict__ fst,
               const float4* __restrict__ v4, const float4* __restrict__ vs4, const float* __restrict__ fixed,
               const float* __restrict__ m_inv, const float* __restrict__ m_inv_st, float mu, int pcap,
               const int32_t* __restrict__ pair_cnt, const float4* __restrict__ pair_rec, int bucket,
               int32_t* __restrict__ bcnt, float4* __restrict__ brec, float4* __restrict__ self_acc,
               int32_t* status) {
  const int nqt = 2 * nv + nvs;
  const int qg = xcd_block() * (int)blockDim.x + (int)threadIdx.x;
  if (qg >= nqt) return;
  const int set = qg < nv ? 0 : (qg < 2 * nv ? 1 : 2);
  const int qv = set == 0 ? qg : (set == 1 ? qg - nv : qg - 2 * nv);
  const int32_t* F = set == 0 ? fst : fdy;
  int n = pair_cnt[qg];
  if (n > pcap) n = pcap;
  float3 acc = zero3();
  int own = 0;
  bool over = false;
  for (int j = 0; j < n; ++j) {
    const size_t col = (size_t)j * 4u * (size_t)nqt + (size_t)qg;
    const float4 r0 = pair_rec[col], r1 = pair_rec[col + (size_t)nqt], r2 = pair_rec[col + 2 * (size_t)nqt],
                 r3 = pair_rec[col + 3 * (size_t)nqt];
    const int fid = __float_as_int(r1.w), dtype = __float_as_int(r2.w);
    const float3 g[4] = {xyz(r0), xyz(r1), xyz(r2), xyz(r3)};
    const int3 fv = make_int3(F[3 * fid], F[3 * fid + 1], F[3 * fid + 2]);
    VelContact c;
    if (!vel_pair(set, qv, fv, dtype, g, r0.w, v4, vs4, fixed, m_inv, m_inv_st, nvs, mu, c)) continue;
    if (set == 0) {
      acc = add3(acc, c.dn[0]);
      acc = add3(acc, c.dt[0]);
      own += 2;
      continue;
    }
    const uint32_t o0 = ((uint32_t)(qg - nv) * (uint32_t)pcap + (uint32_t)j) * 8u;
#pragma unroll
    for (int ph = 0; ph < 2; ++ph)
#pragma unroll
      for (int sl = 0; sl < 4; ++sl) {
        if (!(c.mask & (1u << sl))) continue;
        const float3 d = ph == 0 ? c.dn[sl] : c.dt[sl];
        const int t = c.vid[sl];
        const int at = atomicAdd(&bcnt[t], 1);
        if (at < bucket)
          brec[(size_t)t * bucket + at] = make_float4(d.x, d.y, d.z, __uint_as_float(o0 + 4u * ph + (uint32_t)sl));
        else
          over = true;
      }
  }
  if (qg < nv) self_acc[qg] = make_float4(acc.x, acc.y, acc.z, (float)own);
  if (over && status) atomicOr(status, kStatusRecOverflow);
}

// update_dv (XPBD.py:397-404) over the summed deltas, then (x4 non-null) update_x.
// bcnt null: no deltas (dv = 0, nc = 0).
__global__ void __launch_bounds__(256)
k_vel_apply(int nv, int bucket, int32_t* __restrict__ bcnt, const float4* __restrict__ brec,
            const float4* __restrict__ self_acc, const float* __restrict__ fixed, const float* __restrict__ m_inv,
            float4* __restrict__ v4, float4* __restrict__ x4, float dt) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  float3 dv = zero3();
  float nc = 0.f;
  if (bcnt) {
    const float4 sa = self_acc[v];
    int n = bcnt[v];
    if (n != 0) bcnt[v] = 0;               // left at zero for the next pass
    if (n > bucket) n = bucket;
    nc = sa.w + (float)n;
    dv = bucket_sum(make_float3(sa.x, sa.y, sa.z), brec + (size_t)v * bucket, n);
  }
  if (nc > 0.f) dv = make_float3(dv.x / nc, dv.y / nc, dv.z / nc);
  float4 vv = v4[v];
  if (m_inv[v] > 0.f) {
    const float f = fixed[v];
    vv = make_float4(vv.x + f * dv.x, vv.y + f * dv.y, vv.z + f * dv.z, vv.w);
    v4[v] = vv;
  }
  if (x4) {
    const float4 x = x4[v];
    x4[v] = make_float4(x.x + dt * vv.x, x.y + dt * vv.y, x.z + dt * vv.z, x.w);
  }
}

__global__ void k_compute_v(int nv, const float4* __restrict__ x4, float4* __restrict__ v4,
                            const float4* __restrict__ y4, const float* __restrict__ fixed, float damping, float dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const float4 x = x4[i], y = y4[i];
  const float f = fixed[i];
  const float af = (1.0f - damping) * f;
  v4[i] = make_float4((af * (y.x - x.x)) / dt, (af * (y.y - x.y)) / dt, (af * (y.z - x.z)) / dt, f);
}

__global__ void k_update_x(int nv, float4* __restrict__ x4, const float4* __restrict__ v4, float dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const float4 x = x4[i], v = v4[i];
  x4[i] = make_float4(x.x + dt * v.x, x.y + dt * v.y, x.z + dt * v.z, x.w);
}

__global__ void k_zero_counts(int n, int32_t* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

}  // namespace slx

using namespace slx;

extern "C" int slx_compute_velocity(int nv, const float* x4, float* v4, const float* y4, const float* fixed,
                                    float damping, float dt, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0, "slx_compute_velocity: bad size");
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(x4 && v4 && y4 && fixed, "slx_compute_velocity: null pointer");
  SLX_LAUNCH("compute_velocity", k_compute_v, dim3(blocks_for(nv, 256)), dim3(256), 0, (hipStream_t)stream, nv,
             (const float4*)x4, (float4*)v4, (const float4*)y4, fixed, damping, dt);
  return check_launch("slx_compute_velocity");
}

extern "C" int slx_update_x(int nv, float* x4, const float* v4, float dt, slx_stream_t stream) {
  SLX_REQUIRE(nv >= 0, "slx_update_x: bad size");
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(x4 && v4, "slx_update_x: null pointer");
  SLX_LAUNCH("update_x", k_update_x, dim3(blocks_for(nv, 256)), dim3(256), 0, (hipStream_t)stream, nv,
             (float4*)x4, (const float4*)v4, dt);
  return check_launch("slx_update_x");
}

namespace slx {
int velocity_pass_impl(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st, float* v4,
                       const float* vs4, const float* fixed, const float* m_inv, const float* m_inv_st, float mu,
                       int pair_cap, const int32_t* pair_cnt, const float* pair_rec, const slx_edge_sets* ee,
                       float* x4, float dt, int rec_cap, int32_t* status, void* ws, size_t ws_bytes, hipStream_t st,
                       bool bcnt_zeroed) {
  SLX_REQUIRE(nv >= 0 && nvs >= 0 && pair_cap >= 0 && rec_cap >= 1, "slx_velocity_pass: bad sizes");
  if (nv == 0) return SLX_OK;
  SLX_REQUIRE(v4 && fixed && m_inv, "slx_velocity_pass: null pointer");
  if (ee) {
    SLX_REQUIRE(pair_cnt, "slx_velocity_pass_ee: edge pairs come with the vertex-triangle pair records");
    int rc = ee_check(ee, true);
    if (rc) return rc;
  }
  const bool pairs = pair_cnt != nullptr;
  if (!pairs) {
    SLX_LAUNCH("vel_apply", k_vel_apply, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, 1, nullptr, nullptr,
               nullptr, fixed, m_inv, (float4*)v4, (float4*)x4, dt);
    return check_launch("slx_velocity_pass");
  }
  SLX_REQUIRE(pair_rec && faces_dy && ws, "slx_velocity_pass: pair records need pair_rec, faces and a workspace");
  SLX_REQUIRE(nvs == 0 || faces_st, "slx_velocity_pass: static faces missing");
  SLX_REQUIRE((long long)(nv + nvs) * (long long)pair_cap * 8ll < (1ll << 32),
              "slx_velocity_pass: too many queries x pairs for 32-bit order keys");
  if (ws_bytes < contact_ws_layout(nv, rec_cap, nullptr, nullptr))
    return fail(SLX_EWS, "slx_velocity_pass: workspace too small");
  ContactWs w;
  contact_ws_layout(nv, rec_cap, (char*)ws, &w);
  const int nqt = 2 * nv + nvs;
  if (!bcnt_zeroed) SLX_LAUNCH("fill", k_zero_counts, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, w.bcnt);
  SLX_LAUNCH("vel_contacts", k_vel_contacts, dim3(blocks_for(nqt, 256)), dim3(256), 0, st, nv, nvs, faces_dy,
             faces_st, (const float4*)v4, (const float4*)vs4, fixed, m_inv, m_inv_st, mu, pair_cap, pair_cnt,
             (const float4*)pair_rec, rec_cap, w.bcnt, w.brec, w.self_acc, status);
  if (ee) {
    int rc = ee_velocity_launch(nv, ee, (const float4*)v4, (const float4*)vs4, fixed, m_inv, w, rec_cap, status, st);
    if (rc) return rc;
  }
  SLX_LAUNCH("vel_apply", k_vel_apply, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, rec_cap, w.bcnt, w.brec,
             w.self_acc, fixed, m_inv, (float4*)v4, (float4*)x4, dt);
  return check_launch("slx_velocity_pass");
}
}  // namespace slx

extern "C" int slx_velocity_pass_ee(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st, float* v4,
                                    const float* vs4, const float* fixed, const float* m_inv, const float* m_inv_st,
                                    float mu, int pair_cap, const int32_t* pair_cnt, const float* pair_rec,
                                    const slx_edge_sets* ee, float* x4, float dt, int rec_cap, int32_t* status,
                                    void* ws, size_t ws_bytes, slx_stream_t stream) {
  return velocity_pass_impl(nv, nvs, faces_dy, faces_st, v4, vs4, fixed, m_inv, m_inv_st, mu, pair_cap, pair_cnt,
                            pair_rec, ee, x4, dt, rec_cap, status, ws, ws_bytes, (hipStream_t)stream, false);
}

extern "C" int slx_velocity_pass(int nv, int nvs, const int32_t* faces_dy, const int32_t* faces_st, float* v4,
                                 const float* vs4, const float* fixed, const float* m_inv, const float* m_inv_st,
                                 float mu, int pair_cap, const int32_t* pair_cnt, const float* pair_rec, float* x4,
                                 float dt, int rec_cap, int32_t* status, void* ws, size_t ws_bytes,
                                 slx_stream_t stream) {
  return slx_velocity_pass_ee(nv, nvs, faces_dy, faces_st, v4, vs4, fixed, m_inv, m_inv_st, mu, pair_cap, pair_cnt,
                              pair_rec, nullptr, x4, dt, rec_cap, status, ws, ws_bytes, stream);
}

// ===================== sewing constraints (XPBD.py:418-441) =====================
//
// A Jacobi pass of their own (dx/nc zeroed, the pairs, update_dx): each sewn
// vertex sums the deltas of its pairs in pair order (the serial scatter's order,
// from a host-built list), then y += f dx / nc.  A coincident pair adds no delta
// (the reference would normalise a zero vector) and still counts.
namespace slx {

__global__ void __launch_bounds__(256)
k_sew_gather(int n_sewn, const int32_t* __restrict__ sew_verts, const int32_t* __restrict__ sew_off,
             const int32_t* __restrict__ sew_list, const int2* __restrict__ pairs, const float4* __restrict__ y4,
             float k, float4* __restrict__ ws) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_sewn) return;
  float3 dx = zero3();
  float nc = 0.f;
  for (int e = sew_off[i]; e < sew_off[i + 1]; ++e) {
    const int w = sew_list[e], p = w >> 1, side = w & 1;
    const int2 pr = pairs[p];
    const float4 y1 = y4[pr.x], y2 = y4[pr.y];
    const float3 d = sub3(xyz(y2), xyz(y1));
    const float C = sqrtf(dot3(d, d));
    if (C > 0.f) {
      const float inv = 1.0f / C;
      const float3 n = make_float3(-(inv * d.x), -(inv * d.y), -(inv * d.z));
      const float fw1 = y1.w, fw2 = y2.w;                    // f * m_inv of each point
      const float schur = fw1 + fw2;
      const float ld = (k * C) / ((k * schur) + 1.0f);
      if (side == 0) dx = sub3(dx, mul3(fw1 * ld, n));
      else dx = add3(dx, mul3(fw2 * ld, n));
    }
    nc = nc + 1.0f;
  }
  ws[i] = make_float4(dx.x, dx.y, dx.z, nc);
}

__global__ void __launch_bounds__(256)
k_sew_apply(int n_sewn, const int32_t* __restrict__ sew_verts, const float4* __restrict__ ws,
            const float* __restrict__ fixed, float4* __restrict__ y4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_sewn) return;
  const int v = sew_verts[i];
  const float4 a = ws[i];
  if (!(a.w > 0.f)) return;
  const float f = fixed[v];
  const float4 y = y4[v];
  y4[v] = make_float4(y.x + f * (a.x / a.w), y.y + f * (a.y / a.w), y.z + f * (a.z / a.w), y.w);
}

}  // namespace slx

extern "C" int slx_sewing_pass(int n_sewn, const int32_t* sew_verts, const int32_t* sew_off, const int32_t* sew_list,
                               const int32_t* pairs, float* y4, const float* fixed, float k, float* ws,
                               slx_stream_t stream) {
  SLX_REQUIRE(n_sewn >= 0, "slx_sewing_pass: bad size");
  if (n_sewn == 0) return SLX_OK;
  SLX_REQUIRE(sew_verts && sew_off && sew_list && pairs && y4 && fixed && ws, "slx_sewing_pass: null pointer");
  hipStream_t st = (hipStream_t)stream;
  SLX_LAUNCH("sew_gather", k_sew_gather, dim3(blocks_for(n_sewn, 256)), dim3(256), 0, st, n_sewn, sew_verts, sew_off,
             sew_list, (const int2*)pairs, (const float4*)y4, k, (float4*)ws);
  SLX_LAUNCH("sew_apply", k_sew_apply, dim3(blocks_for(n_sewn, 256)), dim3(256), 0, st, n_sewn, sew_verts,
             (const float4*)ws, fixed, (float4*)y4);
  return check_launch("slx_sewing_pass");
}
