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

/* (f, m_inv) -> x4.w = y4.w = f*m_inv, v4.w = f (set_fixed_vertices,
 * XPBD.py:389-395, with the per-vertex products the kernels read). */
int slx_set_fixed(int nv, const float* fixed, const float* m_inv, float* x4, float* v4,
                  float* y4, slx_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif
