/*
 * CPU oracle for the XPBD cloth step of starlab-physics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * test-suite, by the smoke check and by the benchmark's CPU-baseline leg, and
 * never by the product path.  Every function is a serial restatement of the
 * reference's algorithm (file:line citations beside each declaration, paths
 * relative to the reference repository root) written in plain C with IEEE
 * single/double arithmetic and floating-point contraction disabled.
 *
 * Layout follows the reference's Taichi fields: vec3 arrays are AoS
 * float[3*n] / double[3*n]; scalars per vertex are float[n] / double[n].
 * Functions suffixed _f compute in float, _d in double.
 *
 * Parity status of each function (what pins it) is listed in the test-suite
 * headers; all functions here are pinned (no function is "parity unpinned").
 */
#ifndef STARLAB_ORACLE_H
#define STARLAB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Edge-edge sets of one collision pass (XPBD.py:75-89 pair fields, 217-232 broad
   phase, 266-281 and 313-330 loops).  Candidate faces per dynamic edge, row-major
   [ne][cap] with counts; a null candidate array leaves that set out.  Pair
   records [ne][pcap]: pi int[2] (other edge, type), pg R[12] (g0..g3), ps R[1]
   (schur), nullable; pn_* (per edge) count the contacting pairs, uncapped.  The
   velocity pass reads pn_*, pi, pg, ps. */
typedef struct orc_ee {
  int ne;                                      /* dynamic edges (queries) */
  const int* edges_dy;                         /* [ne][2] */
  const int* edges_st;                         /* static edges [][2] */
  const int* fe_dy;                            /* face -> edge ids, [F][3] */
  const int* fe_st;                            /* [Fs][3] */
  int cap;                                     /* candidate list width */
  const int *st_c, *st_n, *dy_c, *dy_n;        /* candidate faces of the static / dynamic trees */
  int pcap;                                    /* pair records kept per edge */
  int *pn_st, *pn_dy;
  int* pi_st;
  void *pg_st, *ps_st;                         /* R arrays */
  int* pi_dy;
  void *pg_dy, *ps_dy;
} orc_ee;

#define ORC_DECL_REAL(R, S)                                                                 \
  /* XPBD.py:115-118 compute_y */                                                            \
  void orc_compute_y_##S(int nv, const R* x, const R* v, const R* fixed, R* y, R dt,        \
                         const R* g);                                                       \
  /* XPBD.py:345-348 + 140-184 + 383-387: zero dx/nc, spring pass, update_dx */              \
  void orc_spring_pass_##S(int nv, int ne, int nb, const int* edges, const R* l0,           \
                           const int* bend, const R* bl0, const R* x, R* y, const R* fixed, \
                           const R* m_inv, R k_s, R k_b, R* dx, R* nc);                      \
  /* Gauss-Seidel iteration: bending Jacobi pass, then stretch colours in order           \
     (colouring of graph_coloring.py:131-201; the sweep is this package's definition) */   \
  void orc_gs_pass_##S(int nv, int nb, const int* bend, const R* bl0, const R* x, R* y,     \
                       const R* fixed, const R* m_inv, R k_b, const int* gs_order,          \
                       const int* edges, const R* l0, const int* color_off, int n_colors,   \
                       R k_s, R* dx, R* nc);                                                \
  /* residual solve (this package's definition; no reference counterpart): one Jacobi    \
     pass of the implicit-Euler stretch problem anchored at ya (stretch constraints of     \
     XPBD.py:146-167 on y); *res = largest step component of the input y */               \
  void orc_implicit_pass_##S(int nv, int ne, const int* edges, const R* l0, const R* ya,   \
                             R* y, const R* fixed, const R* m_inv, R kappa, R* gs, R* ks,  \
                             R* res);                                                       \
  /* XPBD.py:340-343 compute_velocity then XPBD.py:332-337 update_x */                       \
  void orc_update_vx_##S(int nv, R* x, R* v, const R* y, const R* fixed, R damping, R dt);  \
  /* distance.py:4-47 d_type_PT */                                                           \
  int orc_d_type_pt_##S(const R* p, const R* t0, const R* t1, const R* t2);                 \
  /* distance.py:101-116 (squared distances) + gradients of those (own derivation) */       \
  R orc_dist_##S(int dtype, const R* p, const R* t0, const R* t1, const R* t2, R* g12);     \
  /* XPBD.py:236-264 + 383-387 with collision_constraints_x.py:5-310: zero dx/nc,           \
     vertex-triangle constraints over the three candidate sets, update_dx.                   \
     Candidates: vt_st [nv*cap] + counts, vt_dy [nv*cap], tv_st [nvs*cap].                  \
     pair_num_* receive the recorded pair counts (capped at cap). */                        \
  void orc_collision_pass_##S(int nv, int nvs, const int* faces_dy, const int* faces_st,    \
                              R* y, const R* xs, const R* fixed, const R* m_inv, R dhat,     \
                              R k_col, int cap, const int* vt_st_c, const int* vt_st_n,     \
                              const int* vt_dy_c, const int* vt_dy_n, const int* tv_st_c,   \
                              const int* tv_st_n, R* dx, R* nc, int* pair_num_vt_st,        \
                              int* pair_num_vt_dy, int* pair_num_tv_st);                    \
  /* the same pass, also storing the pair records of each set k (0 vt_st, 1 vt_dy,           \
     2 tv_st) in the reference's layout [query][cap]: pi_k int[2] (face, type),              \
     pg_k R[12] (g0..g3), ps_k R[1] (schur); nullable */                                      \
  void orc_collision_pass_rec_##S(int nv, int nvs, const int* faces_dy, const int* faces_st,\
                                  R* y, const R* xs, const R* fixed, const R* m_inv, R dhat, \
                                  R k_col, int cap, const int* vt_st_c, const int* vt_st_n, \
                                  const int* vt_dy_c, const int* vt_dy_n,                   \
                                  const int* tv_st_c, const int* tv_st_n, R* dx, R* nc,     \
                                  int* pair_num_vt_st, int* pair_num_vt_dy,                 \
                                  int* pair_num_tv_st, int* pi0, R* pg0, R* ps0, int* pi1,  \
                                  R* pg1, R* ps1, int* pi2, R* pg2, R* ps2);                \
  /* XPBD.py:418-441 sewing constraints as a Jacobi pass (zero dx/nc, pairs in               \
     order, update_dx); pairs int[2 np] */                                                   \
  void orc_sewing_pass_##S(int nv, int np, const int* pairs, R* y, const R* fixed,          \
                           const R* m_inv, R k, R* dx, R* nc);                               \
  /* XPBD.py:340-343 compute_velocity */                                                     \
  void orc_compute_velocity_##S(int nv, const R* x, R* v, const R* y, const R* fixed,       \
                                R damping, R dt);                                           \
  /* XPBD.py:332-337 update_x */                                                             \
  void orc_update_x_##S(int nv, R* x, const R* v, R dt);                                    \
  /* XPBD.py:374-381 solve_constraints_v with 283-311 and                                   \
     collision_constraints_v.py:5-570 over the recorded pairs, then update_dv               \
     (XPBD.py:397-404); v updated in place, dv/nc are scratch outputs */                     \
  void orc_velocity_pass_##S(int nv, int nvs, const int* faces_dy, const int* faces_st,     \
                             R* v, const R* vs, const R* fixed, const R* m_inv,             \
                             const R* m_inv_st, R mu, int cap, const int* pn0,              \
                             const int* pi0, const R* pg0, const R* ps0, const int* pn1,    \
                             const int* pi1, const R* pg1, const R* ps1, const int* pn2,    \
                             const int* pi2, const R* pg2, const R* ps2, R* dv, R* nc);     \
  /* distance.py:49-99 d_type_EE (8 = both segment interiors) */                             \
  int orc_d_type_ee_##S(const R* x0, const R* x1, const R* x2, const R* x3);                \
  /* distance.py:118-124 d_EE (squared line-line distance) + gradient (own derivation) */    \
  R orc_dist_ee_##S(const R* x0, const R* x1, const R* x2, const R* x3, R* g12);            \
  /* orc_collision_pass_rec with the edge-edge loops after the vertex-triangle ones          \
     (XPBD.py:266-281, collision_constraints_x.py:312-507), then update_dx */               \
  void orc_collision_pass_ee_##S(int nv, int nvs, const int* faces_dy, const int* faces_st, \
                                 R* y, const R* xs, const R* fixed, const R* m_inv, R dhat,  \
                                 R k_col, int cap, const int* vt_st_c, const int* vt_st_n,  \
                                 const int* vt_dy_c, const int* vt_dy_n,                    \
                                 const int* tv_st_c, const int* tv_st_n, R* dx, R* nc,      \
                                 int* pair_num_vt_st, int* pair_num_vt_dy,                  \
                                 int* pair_num_tv_st, int* pi0, R* pg0, R* ps0, int* pi1,   \
                                 R* pg1, R* ps1, int* pi2, R* pg2, R* ps2, const orc_ee* ee); \
  /* orc_velocity_pass with the edge-edge pairs after the vertex-triangle ones              \
     (XPBD.py:313-330, collision_constraints_v.py:573-878), then update_dv */              \
  void orc_velocity_pass_ee_##S(int nv, int nvs, const int* faces_dy, const int* faces_st,  \
                                R* v, const R* vs, const R* fixed, const R* m_inv,          \
                                const R* m_inv_st, R mu, int cap, const int* pn0,           \
                                const int* pi0, const R* pg0, const R* ps0, const int* pn1, \
                                const int* pi1, const R* pg1, const R* ps1, const int* pn2, \
                                const int* pi2, const R* pg2, const R* ps2, R* dv, R* nc,  \
                                const orc_ee* ee);

ORC_DECL_REAL(float, f)
ORC_DECL_REAL(double, d)

/* ---- LBVH (framework/collision/lbvh_cell.py), float + int32 only ---- */

/* lbvh_cell.py:85-92 */
uint32_t orc_expand_bits(uint32_t v);
/* lbvh_cell.py:94-102 (x in the lowest bit) */
int32_t orc_morton3d(float x, float y, float z);
/* lbvh_cell.py:122-152 assign_cell_morton over the 32^3 grid of [lo, hi] */
void orc_cell_morton(const float* lo, const float* hi, int32_t* codes, float* centers,
                     float* cell_size);
/* lbvh_cell.py:524-562 serial LSD radix sort (digit width bits, back-to-front scatter) */
void orc_radix_sort(int n, int32_t* keys, int32_t* ids, int bits, int passes);
/* lbvh_cell.py:255-336, 407-430 Karras 2012 over sorted codes.  wide_keys=1 appends
   the index to each code (64-bit keys) so duplicate codes build a valid tree. */
void orc_karras(int n, const int32_t* sorted_codes, int wide_keys, int32_t* child_a,
                int32_t* child_b, int32_t* parent, int32_t* range_l, int32_t* range_r);
/* meshTaichiWrapper.py:350-361 computeAABB_faces */
void orc_face_aabbs(int nf, const int* faces, const float* x, float padding, float* fmin,
                    float* fmax);
/* meshTaichiWrapper.py:334-347 computeAABB */
void orc_mesh_aabb(int nv, const float* x, float padding, float* lo, float* hi);
/* lbvh_cell.py:564-574: truncated cell index, flattened with int32 wrap-around */
int32_t orc_flat_cell_id(const float* pos, const float* cell_size, const float* origin);

/* lbvh_cell.py:637-653 LBVH_CELL.build over a static Karras topology.
   Node arrays have 2*32768-1 entries, aabbs are float[3*nodes]. */
int orc_cell_build(int nf, const float* fmin, const float* fmax, const float* lo,
                   const float* hi, const int32_t* child_a, const int32_t* child_b,
                   const int32_t* parent, float* cell_size, int32_t* face_cell,
                   int32_t* prefix, int32_t* sorted_faces, int32_t* num_faces,
                   int32_t* leaf_l, int32_t* leaf_r, float* nmin, float* nmax);

/* Per-triangle LBVH: Morton codes of face-box centres normalised by the cube
   [lo, lo + e] (e = [lo,hi]'s largest extent, the same scale on every axis),
   stable sort, Karras with index tie-break, leaves = sorted faces, refit. */
void orc_tri_build(int nf, const float* fmin, const float* fmax, const float* lo,
                   const float* hi, int32_t* codes, int32_t* sorted_faces, int32_t* child_a,
                   int32_t* child_b, int32_t* parent, float* nmin, float* nmax);

/* Refit of a triangle LBVH over new face boxes, topology kept. */
void orc_tri_refit(int nf, const float* fmin, const float* fmax, const int32_t* sorted_faces,
                   const int32_t* child_a, const int32_t* child_b, float* nmin, float* nmax);

/* lbvh_cell.py:655-714 traverse_cell_bvh_single for a batch of boxes.  mode 0 =
   cell tree (leaf ranges into sorted_faces), mode 1 = triangle tree (one face per
   leaf, leaf k holds sorted_faces[k]).  Returns the number of queries whose list
   overflowed cap (entries past cap are dropped) or whose stack overflowed. */
int orc_traverse(int mode, int nq, const float* qmin, const float* qmax, int n_leaves,
                 const int32_t* child_a, const int32_t* child_b, const int32_t* num_faces,
                 const int32_t* leaf_l, const int32_t* leaf_r, const int32_t* sorted_faces,
                 const float* nmin, const float* nmax, const float* fmin, const float* fmax,
                 const float* cell_size, const float* origin, int cap, int32_t* cand,
                 int32_t* cand_n);

/* graph_coloring.py:131-149 greedy colouring + 172-201 stable counting sort */
int orc_greedy_coloring(int nv, int ne, const int* edges, int max_colors, int32_t* color,
                        int32_t* order, int32_t* offsets);

#ifdef __cplusplus
}
#endif
#endif
