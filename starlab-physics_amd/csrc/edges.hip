// Edge-edge constraints on gfx950: position-level (after the vertex-triangle
// contacts, before update_dx) and velocity-level (after the vertex-triangle
// pairs, before update_dv), with deterministic accumulation.
//
// Reference (all of it commented out of the reference's step, kept callable
// there): framework/physics/XPBD.py:75-89 (pair fields), 217-232 (edge broad
// phase), 266-281 and 313-330 (loops); framework/physics/
// collision_constraints_x.py:312-507 (__ee_dy, __ee_st), collision_constraints_v.py
// 573-878 (__ee_st, __ee_dy); framework/collision/distance.py:49-124 (d_type_EE,
// d_EE; the gradient is derived here).
//
// The edge sets come after the vertex-triangle sets in the reference's serial
// order, so a vertex's total is (its vertex-triangle sum) + its edge-edge deltas
// in edge order.  The vertex-triangle records are first folded into each
// vertex's running sum (k_bucket_fold, same summation order), the buckets are
// emptied, and the edge queries put their deltas in the buckets with keys in
// their own serial order:
//   set 0 (static):  ((e * cap + j) * 3 + side) * 2 + slot     (slots x0, x1)
//   set 1 (dynamic): 6 ne cap + ((e * cap + j) * 3 + side) * 4 + slot
// and for the velocity pass, over pair records j:
//   set 0: (e * pair_cap + j) * 2 + slot,  set 1: 2 ne pair_cap + (e * pair_cap + j) * 4 + slot.
#include "slx_common.h"

namespace slx {

// d_type_EE, distance.py:49-99 (op order as the serial restatement)
__device__ int d_type_ee(float3 x0, float3 x1, float3 x2, float3 x3) {
  const float3 u = sub3(x1, x0), v = sub3(x3, x2), w = sub3(x0, x2);
  const float a = dot3(u, u), b = dot3(u, v), c = dot3(v, v), d = dot3(u, w), e = dot3(v, w);
  const float D = a * c - b * b;
  float tD = D, tN;
  int ty = 8;
  const float sN = b * e - c * d;
  if (sN <= 0.f) {
    tN = e;
    tD = c;
    ty = 2;
  } else if (sN >= D) {
    tN = e + b;
    tD = c;
    ty = 5;
  } else {
    tN = a * e - b * d;
    const float3 uv = cross3(u, v);
    if (tN > 0.f && tN < tD && (dot3(uv, w) < 1e-4f || dot3(uv, uv) < (1.0e-20f * a) * c)) {
      if (sN < D / 2.0f) {
        tN = e;
        tD = c;
        ty = 2;
      } else {
        tN = e + b;
        tD = c;
        ty = 5;
      }
    }
  }
  if (tN <= 0.f) {
    if (-d <= 0.f) ty = 0;
    else if (-d >= a) ty = 3;
    else ty = 6;
  } else if (tN >= tD) {
    if ((-d + b) <= 0.f) ty = 1;
    else if ((-d + b) >= a) ty = 4;
    else ty = 7;
  }
  return ty;
}

// d_EE, distance.py:118-124: d = s^2/N, u = x1-x0, v = x3-x2, w = x2-x0, b = u x v,
// s = w.b, N = b.b; gradient g_k = (2s/N) ds_k - (s^2/N^2) dN_k with
// ds = (-b - v x w, v x w, b - w x u, w x u), dN = 2 (-(v x b), v x b, -(b x u), b x u).
__device__ __forceinline__ float dist_ee(float3 x0, float3 x1, float3 x2, float3 x3, float3 (&g)[4]) {
  const float3 u = sub3(x1, x0), v = sub3(x3, x2), w = sub3(x2, x0);
  const float3 b = cross3(u, v);
  const float s = dot3(w, b), N = dot3(b, b);
  const float d = (s * s) / N;
  const float3 vw = cross3(v, w), wu = cross3(w, u), vb = cross3(v, b), bu = cross3(b, u);
  const float p = (2.0f * s) / N, q = (s * s) / (N * N);
  g[0] = make_float3(p * (-b.x - vw.x) - q * (2.0f * -vb.x), p * (-b.y - vw.y) - q * (2.0f * -vb.y),
                     p * (-b.z - vw.z) - q * (2.0f * -vb.z));
  g[1] = make_float3(p * vw.x - q * (2.0f * vb.x), p * vw.y - q * (2.0f * vb.y), p * vw.z - q * (2.0f * vb.z));
  g[2] = make_float3(p * (b.x - wu.x) - q * (2.0f * -bu.x), p * (b.y - wu.y) - q * (2.0f * -bu.y),
                     p * (b.z - wu.z) - q * (2.0f * -bu.z));
  g[3] = make_float3(p * wu.x - q * (2.0f * bu.x), p * wu.y - q * (2.0f * bu.y), p * wu.z - q * (2.0f * bu.z));
  return d;
}

struct EdgeArgs {
  int ne, cap, pcap;
  const int32_t *edges_dy, *edges_st, *fe_dy, *fe_st;
  const int32_t *c_st, *n_st, *c_dy, *n_dy;
  int32_t* pair_num;
  int32_t* pair_cnt;
  float4* pair_rec;
};

// A vertex's bucket records so far folded into its running sum (in key order,
// as the apply kernels would add them) and count; the bucket is emptied.
__global__ void __launch_bounds__(256)
k_bucket_fold(int nv, int bucket, int32_t* __restrict__ bcnt, const float4* __restrict__ brec,
              float4* __restrict__ self_acc) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const float4 sa = self_acc[v];
  int n = bcnt[v];
  if (n > bucket) n = bucket;
  const float3 acc = bucket_sum(make_float3(sa.x, sa.y, sa.z), brec + (size_t)v * bucket, n);
  self_acc[v] = make_float4(acc.x, acc.y, acc.z, sa.w + (float)n);
  bcnt[v] = 0;
}

__device__ __forceinline__ void push_rec(int t, float3 d, uint32_t key, int bucket, int32_t* __restrict__ bcnt,
                                         float4* __restrict__ brec, bool& over) {
  const int at = atomicAdd(&bcnt[t], 1);
  if (at < bucket) brec[(size_t)t * bucket + at] = make_float4(d.x, d.y, d.z, __uint_as_float(key));
  else over = true;
}

// One thread per (set, edge): the sides of its candidate faces in order
// (collision_constraints_x.py:312-507).  (Threads take the edges in index order:
// the candidate lists are read coalesced.)
__global__ void __launch_bounds__(256)
k_ee_contacts(EdgeArgs A, const float4* __restrict__ y4, const float4* __restrict__ xs4,
              const float* __restrict__ fixed, const float* __restrict__ m_inv, float dhat, float k, int bucket,
              int32_t* __restrict__ bcnt, float4* __restrict__ brec, int32_t* status) {
  const int nq = 2 * A.ne;
  const int qg = blockIdx.x * blockDim.x + threadIdx.x;
  if (qg >= nq) return;
  const int set = qg >= A.ne ? 1 : 0;
  const int e = qg - set * A.ne;
  const int32_t* C = set ? A.c_dy : A.c_st;
  int pairs = 0, stored = 0;
  bool over = false, pover = false;
  if (C) {
    const int32_t* N = set ? A.n_dy : A.n_st;
    const int32_t* fe = set ? A.fe_dy : A.fe_st;
    const int32_t* E2 = set ? A.edges_dy : A.edges_st;
    const float4* P2 = set ? y4 : xs4;
    int n = N[e];
    if (n > A.cap) n = A.cap;
    const int v0 = A.edges_dy[2 * e], v1 = A.edges_dy[2 * e + 1];
    const float3 x0 = xyz(y4[v0]), x1 = xyz(y4[v1]);
    const float3 x01 = sub3(x0, x1);
    const uint32_t base = set ? 6u * (uint32_t)A.ne * (uint32_t)A.cap : 0u;
    const uint32_t nslot = set ? 4u : 2u;
    for (int j = 0; j < n; ++j) {
      const int f = C[(size_t)j * A.ne + e];
      const int3 sides = make_int3(fe[3 * f], fe[3 * f + 1], fe[3 * f + 2]);
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        const int e2 = kk == 0 ? sides.x : (kk == 1 ? sides.y : sides.z);
        const int v2 = E2[2 * e2], v3 = E2[2 * e2 + 1];
        if (set == 1 && (v0 == v2 || v0 == v3 || v1 == v2 || v1 == v3)) continue;
        const float3 x2 = xyz(P2[v2]), x3 = xyz(P2[v3]);
        const int dtype = d_type_ee(x0, x1, x2, x3);
        if (dtype != 8) continue;
        const float3 cr = cross3(x01, sub3(x2, x3));
        if (!(sqrtf(dot3(cr, cr)) > 1e-3f)) continue;
        float3 g[4];
        const float d = dist_ee(x0, x1, x2, x3, g);
        if (!(d < dhat)) continue;
        const int vid[4] = {v0, v1, v2, v3};
        const int nmov = set ? 4 : 2;
        float fw[4];
        float schur = 0.f;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          if (s >= nmov) break;
          fw[s] = fixed[vid[s]] * m_inv[vid[s]];
          const float term = fw[s] * dot3(g[s], g[s]);
          schur = s == 0 ? term : schur + term;
        }
        const float ld = (k * (dhat - d)) / ((k * schur) + 1.0f);
        const uint32_t o0 = base + (((uint32_t)e * (uint32_t)A.cap + (uint32_t)j) * 3u + (uint32_t)kk) * nslot;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          if (s >= nmov) break;
          push_rec(vid[s], mul3(fw[s] * ld, g[s]), o0 + (uint32_t)s, bucket, bcnt, brec, over);
        }
        if (A.pair_rec) {
          if (stored < A.pcap) {
            const size_t col = (size_t)stored * 4u * (size_t)nq + (size_t)qg;
            A.pair_rec[col] = make_float4(g[0].x, g[0].y, g[0].z, schur);
            A.pair_rec[col + (size_t)nq] = make_float4(g[1].x, g[1].y, g[1].z, __int_as_float(e2));
            A.pair_rec[col + 2 * (size_t)nq] = make_float4(g[2].x, g[2].y, g[2].z, __int_as_float(dtype));
            A.pair_rec[col + 3 * (size_t)nq] = make_float4(g[3].x, g[3].y, g[3].z, 0.f);
            ++stored;
          } else {
            pover = true;
          }
        }
        ++pairs;
      }
    }
  }
  if (A.pair_num) A.pair_num[qg] = pairs;
  if (A.pair_rec) A.pair_cnt[qg] = stored;
  if (over && status) atomicOr(status, kStatusRecOverflow);
  if (pover && status) atomicOr(status, kStatusPairOverflow);
}

// One thread per (set, edge): its recorded pairs in order
// (collision_constraints_v.py:573-878, normal part only).
__global__ void __launch_bounds__(256)
k_vel_ee_contacts(EdgeArgs A, const float4* __restrict__ v4, const float4* __restrict__ vs4,
                  const float* __restrict__ fixed, const float* __restrict__ m_inv, int bucket,
                  int32_t* __restrict__ bcnt, float4* __restrict__ brec, int32_t* status) {
  const int nq = 2 * A.ne;
  const int qg = blockIdx.x * blockDim.x + threadIdx.x;
  if (qg >= nq) return;
  const int set = qg >= A.ne ? 1 : 0;
  const int e = qg - set * A.ne;
  int n = A.pair_cnt[qg];
  if (n > A.pcap) n = A.pcap;
  if (n <= 0) return;
  const int v0 = A.edges_dy[2 * e], v1 = A.edges_dy[2 * e + 1];
  const float3 u0 = xyz(v4[v0]), u1 = xyz(v4[v1]);
  const uint32_t base = set ? 2u * (uint32_t)A.ne * (uint32_t)A.pcap : 0u;
  const uint32_t nslot = set ? 4u : 2u;
  bool over = false;
  for (int j = 0; j < n; ++j) {
    const size_t col = (size_t)j * 4u * (size_t)nq + (size_t)qg;
    const float4 r0 = A.pair_rec[col], r1 = A.pair_rec[col + (size_t)nq], r2 = A.pair_rec[col + 2 * (size_t)nq],
                 r3 = A.pair_rec[col + 3 * (size_t)nq];
    const int e2 = __float_as_int(r1.w);
    const float3 g0 = xyz(r0), g1 = xyz(r1), g2 = xyz(r2), g3 = xyz(r3);
    const float schur = r0.w;
    const uint32_t o0 = base + ((uint32_t)e * (uint32_t)A.pcap + (uint32_t)j) * nslot;
    if (set == 0) {
      // __ee_st: Cv = g0.v0 + g1.v1 + g3.vs3 (g2 does not enter, collision_constraints_v.py:782)
      const int v3 = A.edges_st[2 * e2 + 1];
      const float3 u3 = vs4 ? xyz(vs4[v3]) : make_float3(0.f, 0.f, 0.f);
      const float Cv = (dot3(g0, u0) + dot3(g1, u1)) + dot3(g3, u3);
      if (!(Cv < 0.f)) continue;
      const float ld = -Cv / schur;
      push_rec(v0, mul3((fixed[v0] * m_inv[v0]) * ld, g0), o0, bucket, bcnt, brec, over);
      push_rec(v1, mul3((fixed[v1] * m_inv[v1]) * ld, g1), o0 + 1u, bucket, bcnt, brec, over);
    } else {
      // __ee_dy: dvn = sum g.v; dv -= f w g (dvn / schur) on all four points
      const int v2 = A.edges_dy[2 * e2], v3 = A.edges_dy[2 * e2 + 1];
      const float dvn = ((dot3(g0, u0) + dot3(g1, u1)) + dot3(g2, xyz(v4[v2]))) + dot3(g3, xyz(v4[v3]));
      if (!(dvn < 0.f)) continue;
      const float ld = dvn / schur;
      const int vid[4] = {v0, v1, v2, v3};
      const float3 g[4] = {g0, g1, g2, g3};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const float a = fixed[vid[s]] * m_inv[vid[s]];
        const float3 d = make_float3(-((a * g[s].x) * ld), -((a * g[s].y) * ld), -((a * g[s].z) * ld));
        push_rec(vid[s], d, o0 + (uint32_t)s, bucket, bcnt, brec, over);
      }
    }
  }
  if (over && status) atomicOr(status, kStatusRecOverflow);
}

static EdgeArgs edge_args(const slx_edge_sets* ee) {
  EdgeArgs A;
  A.ne = ee->ne;
  A.cap = ee->cap;
  A.pcap = ee->pair_cap;
  A.edges_dy = ee->edges_dy;
  A.edges_st = ee->edges_st;
  A.fe_dy = ee->fe_dy;
  A.fe_st = ee->fe_st;
  A.c_st = ee->c_st;
  A.n_st = ee->n_st;
  A.c_dy = ee->c_dy;
  A.n_dy = ee->n_dy;
  A.pair_num = ee->pair_num;
  A.pair_cnt = ee->pair_cnt;
  A.pair_rec = (float4*)ee->pair_rec;
  return A;
}

int ee_check(const slx_edge_sets* ee, bool velocity) {
  SLX_REQUIRE(ee->ne >= 0 && ee->cap >= 0 && ee->pair_cap >= 0, "edge sets: bad sizes");
  if (ee->ne == 0) return SLX_OK;
  SLX_REQUIRE(ee->edges_dy, "edge sets: null edges_dy");
  if (velocity) {
    SLX_REQUIRE(ee->pair_cnt && ee->pair_rec, "edge sets: the velocity pass needs pair_cnt and pair_rec");
    SLX_REQUIRE(ee->edges_st || !ee->c_st, "edge sets: null edges_st");
    SLX_REQUIRE((long long)ee->ne * (long long)ee->pair_cap * 6ll < (1ll << 32),
                "edge sets: too many edges x pairs for 32-bit order keys");
    return SLX_OK;
  }
  SLX_REQUIRE(!ee->c_st || (ee->n_st && ee->fe_st && ee->edges_st), "edge sets: static set incomplete");
  SLX_REQUIRE(!ee->c_dy || (ee->n_dy && ee->fe_dy), "edge sets: dynamic set incomplete");
  SLX_REQUIRE(!ee->pair_rec || ee->pair_cnt, "edge sets: pair records need pair_cnt");
  SLX_REQUIRE((long long)ee->ne * (long long)ee->cap * 18ll < (1ll << 32),
              "edge sets: too many edges x candidates for 32-bit order keys");
  return SLX_OK;
}

// Fold the vertex-triangle records, then the edge-edge deltas into the buckets.
int ee_position_launch(int nv, const slx_edge_sets* ee, const float4* y4, const float4* xs4, const float* fixed,
                       const float* m_inv, float dhat, float k, const ContactWs& w, int bucket, int32_t* status,
                       hipStream_t st) {
  SLX_LAUNCH("bucket_fold", k_bucket_fold, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, bucket, w.bcnt, w.brec,
             w.self_acc);
  if (ee->ne == 0) return check_launch("ee_contacts");
  SLX_REQUIRE(!ee->c_st || xs4, "slx_collision_pass_ee: the static set needs static positions");
  SLX_LAUNCH("ee_contacts", k_ee_contacts, dim3(blocks_for(2ll * ee->ne, 256)), dim3(256), 0, st, edge_args(ee), y4,
             xs4, fixed, m_inv, dhat, k, bucket, w.bcnt, w.brec, status);
  return check_launch("ee_contacts");
}

int ee_velocity_launch(int nv, const slx_edge_sets* ee, const float4* v4, const float4* vs4, const float* fixed,
                       const float* m_inv, const ContactWs& w, int bucket, int32_t* status, hipStream_t st) {
  SLX_LAUNCH("bucket_fold", k_bucket_fold, dim3(blocks_for(nv, 256)), dim3(256), 0, st, nv, bucket, w.bcnt, w.brec,
             w.self_acc);
  if (ee->ne == 0) return check_launch("vel_ee_contacts");
  SLX_LAUNCH("vel_ee_contacts", k_vel_ee_contacts, dim3(blocks_for(2ll * ee->ne, 256)), dim3(256), 0, st,
             edge_args(ee), v4, vs4, fixed, m_inv, bucket, w.bcnt, w.brec, status);
  return check_launch("vel_ee_contacts");
}

}  // namespace slx
