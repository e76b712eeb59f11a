/*
 * funlbm.h - C ABI of the MI355X (gfx950) lattice-Boltzmann + immersed-boundary
 * timestep library (libfunlbm_hip.so).
 *
 * Conventions
 *   - Every entry point returns int: 0 = ok, > 0 = a hipError_t, < 0 = an
 *     argument/shape error.  funlbm_last_error() gives a thread-local message.
 *   - All array arguments are device pointers owned by the caller (PyTorch's
 *     caching allocator in the Python host).  Nothing is allocated inside a call,
 *     so every entry point may be recorded into a HIP graph.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).  Calls
 *     are asynchronous with respect to the host.
 *   - Populations are stored structure-of-arrays: f[k][x][y][z], z fastest,
 *     direction order k exactly as the reference lattice tables
 *     (src/funlbm/flow/d3/d3q13.py:7-44, d3q15.py:7-48, d3q19.py:7-60,
 *     d3q27.py:7-77).  Vector fields are [3][x][y][z], scalar fields [x][y][z].
 *     The reference's [X,Y,Z,Q] tensors are the permuted views of these.
 *   - Node counts are the reference's config size + 1 per axis
 *     (src/funlbm/flow/d3/base.py:81-82).
 */
#ifndef FUNLBM_H
#define FUNLBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FUNLBM_ABI_VERSION 4

/* Boundary codes: the values of the reference enum BoundaryCondition
 * (src/funlbm/config/base.py:22-28). */
#define FUNLBM_BC_PERIODICAL 11000
#define FUNLBM_BC_WALL 1200
#define FUNLBM_BC_WALL_WITH_SPEED 1201
#define FUNLBM_BC_NON_EQUILIBRIUM 1400 /* input/output faces, explicit entry points only */

/* Face order of bc[], uw[] and the bits of uzero_mask: the reference's
 * input(x=0), output(x=X), back(y=0), forward(y=Y), bottom(z=0), top(z=Z)
 * (src/funlbm/flow/d3/base.py:189-196). */

/* Error codes (negative; positive values are hipError_t). */
#define FUNLBM_EINVAL -1                /* bad argument or shape */
#define FUNLBM_EUNSUPPORTED_BC -2       /* boundary code this path does not implement */
#define FUNLBM_EUNSUPPORTED_LATTICE -3  /* lattice other than D3Q13/15/19/27 */
#define FUNLBM_ETOOBIG -4               /* size beyond a documented maximum */
#define FUNLBM_EFAILED -5               /* a device-side check failed */

/* Storage of the population state buffers used by the fused path. */
#define FUNLBM_STORE_F32 0  /* f as float */
#define FUNLBM_STORE_F16 1  /* (f - w_k*rho0)*dev_scale as IEEE binary16 */
#define FUNLBM_STORE_BF16 2 /* (f - w_k*rho0)*dev_scale as bfloat16 */
#define FUNLBM_STORE_F64 3  /* f as double, arithmetic in double */

/* Rounding of the F16/BF16 state at each fused step's store. */
#define FUNLBM_ROUND_NEAREST 0    /* round to nearest even */
#define FUNLBM_ROUND_STOCHASTIC 1 /* round up with probability equal to the discarded
                                     fraction of a storage ulp (gfx950's stochastic
                                     conversion), fed by a hash of (seed, node, direction):
                                     unbiased, so increments below half an ulp of the
                                     stored deviation accumulate instead of being lost */

/* One-dimensional kernel of the immersed-boundary weights (per axis, the
 * 3-D weight is the product, normalised over the stencil). */
#define FUNLBM_DELTA_COSINE 0  /* (1 + cos(pi r/2))/4 for |r| <= 2: the reference's
                                  (src/funlbm/lbm/lbm3d.py:46-64) */
#define FUNLBM_DELTA_PESKIN4 1 /* Peskin's 4-point kernel: (3 - 2|r| + sqrt(1 + 4|r| - 4r^2))/8
                                  for |r| <= 1, (5 - 2|r| - sqrt(-7 + 12|r| - 4r^2))/8 for
                                  1 <= |r| <= 2 (extension; no reference counterpart) */

typedef struct funlbm_lattice {
  int32_t q;          /* 13, 15, 19 or 27 */
  int32_t nx, ny, nz; /* node counts */
  int32_t storage;    /* FUNLBM_STORE_* */
  int32_t bc[6];      /* FUNLBM_BC_* per face */
  int32_t uzero_mask; /* bit i: u := 0 on face i after every moment update (0x3f = reference) */
  float tau;          /* BGK relaxation time, 3*mu + 0.5 */
  float rho0;         /* reference density of deviation storage: must be 1.0
                         for F16/BF16 (the deviation base is w_k); ignored otherwise */
  float dev_scale;    /* scale of deviation storage (ignored for F32/F64) */
  float body_force[3];/* uniform force added to FOL everywhere (extension; 0 = none) */
  float uw[6][3];     /* wall velocity of WALL_WITH_SPEED faces */
  int32_t rounding;   /* FUNLBM_ROUND_*; STOCHASTIC needs F16 or BF16 storage */
  uint32_t seed;      /* stochastic rounding: the dither seed of this call's step
                         (a solve passes base + steps_done * 0x9E3779B9, mod 2^32, so
                         every step draws new dithers); funlbm_case_pack takes the
                         base and the batched step adds its own step count */
  int32_t delta;      /* FUNLBM_DELTA_*: the interpolation/spreading kernel of the
                         immersed-boundary entry points */
} funlbm_lattice;

/* ABI version (FUNLBM_ABI_VERSION).  No reference counterpart. */
int funlbm_abi_version(void);

/* Thread-local description of the last error.  No reference counterpart. */
const char* funlbm_last_error(void);

/* Bytes of one fused-state buffer for `lat`.  The nodes are grouped into
 * tiles of T consecutive nodes (T = 4096 from 2^14 nodes up, else 256); a tile
 * holds its q direction planes of T elements one after another, and one slack
 * tile precedes the first tile and one follows the last:
 * (ceil(n / T) + 2) * T * q elements of lat->storage.  Element (k, node j)
 * sits at q T + ((j / T) q + k) T + j % T.  No reference counterpart. */
int funlbm_state_bytes(const funlbm_lattice* lat, int64_t* bytes);

/* ---- reference-structured operations (explicit path, float, in place) ---- */

/* BGK collision, replaces FlowD3.cul_equ (src/funlbm/flow/d3/base.py:150-159):
 * feq = (1 + (u.e)/cs^2 + ((u.e)/cs^2)^2/2 - |u|^2/cs^2/2) * rho*w,
 * f = f - (f - feq)/tau.  feq may be NULL. */
int funlbm_collide(const funlbm_lattice* lat, float* f, const float* u, const float* rho,
                   float* feq, void* stream);

/* Streaming with the six face overrides, replaces FlowD3.f_stream and
 * f_stream_bound* (src/funlbm/flow/d3/base.py:161-323).  Works on the storage
 * format of lat->storage, so it is a bit-exact permutation for PERIODICAL and
 * WALL faces.  rho (float, [x][y][z]) is read only on WALL_WITH_SPEED faces.
 * feq ([q][x][y][z] float, the last collision's equilibrium) is read only on
 * NON_EQUILIBRIUM faces, where every direction becomes feq + (f - feq) of the
 * next plane in (src/funlbm/flow/d3/base.py:214-218,237-241); those faces
 * need float32 storage and nx >= 3. */
int funlbm_stream(const funlbm_lattice* lat, const void* f_src, void* f_dst, const float* rho,
                  const float* feq, void* stream);

/* Moments, replaces FlowD3.update_u_rou + update_u_rou_boundary
 * (src/funlbm/flow/d3/base.py:116-133): rho = sum f, p = rho/3,
 * u = (f.e + FOL/2)/rho, then u := 0 on the faces of uzero_mask.
 * fol may be NULL (zero); body_force is added to fol.  p may be NULL. */
int funlbm_moments(const funlbm_lattice* lat, const float* f, const float* fol, float* u,
                   float* rho, float* p, void* stream);

/* The NON_EQUILIBRIUM overrides of update_u_rou_boundary
 * (src/funlbm/flow/d3/base.py:135-148), run after funlbm_moments: on an
 * input NEQ face u_x = uw * profile[y][z] (double, the duct profile of
 * init_u for u_max = 1, [ny][nz]) with uw = float32(remu / max(rho)) /
 * min(ny, nz) in float32 (remu = float32(Re * mu)), the other components
 * stay zero, and rho = rho of the next plane in; on an output NEQ face
 * rho and u are copied from the next plane in.  work: 1 float of device
 * scratch (the max).  No-op without NEQ faces. */
int funlbm_neq_moments(const funlbm_lattice* lat, float* u, float* rho, const double* profile, float remu,
                       float* work, void* stream);

/* Forcing term, replaces FlowD3.cul_equ2 (src/funlbm/flow/d3/base.py:103-114):
 * f_a += (1 - 1/(2 tau)) w_a sum_d [(e_ad - u_d)/cs^2 + u_d e_ad e_ad/cs^4] FOL_d. */
int funlbm_forcing(const funlbm_lattice* lat, float* f, const float* u, const float* fol,
                   void* stream);

/* ---- fused path ---- */

/* Converts float populations (the reference's f after cul_equ) into a state
 * buffer of lat->storage.  No reference counterpart (storage format). */
int funlbm_pack(const funlbm_lattice* lat, const float* f, void* state, void* stream);

/* out = equilibrium populations feq(rho, u), [q][x][y][z] in lat->storage
 * (planes of n elements, the layout funlbm_pack reads, not the padded state
 * layout), computed in the storage's arithmetic type (rho NULL = 1, u NULL =
 * 0); FlowD3.init's f = rho w is the u = 0 case
 * (src/funlbm/flow/d3/base.py:94-100). */
int funlbm_equilibrium_state(const funlbm_lattice* lat, const float* rho, const float* u,
                             void* state, void* stream);

/* The state the fused path starts from after the reference's init: f = w
 * (rho = 1), the moments of it, then the extra collision of LBMD3._init and
 * the first step's collision with those moments (src/funlbm/flow/d3/base.py:
 * 99-101, src/funlbm/lbm/lbm3d.py:29, src/funlbm/lbm/base.py:161), computed in
 * the storage's arithmetic type. */
int funlbm_init_state(const funlbm_lattice* lat, void* state, void* stream);

/* One fused pass replacing, for every node, the tail of LBMBase.run_step
 * (src/funlbm/lbm/base.py:159-175: f_stream, update_u_rou, cul_equ2,
 * update_u_rou) and the cul_equ that opens the next step (:161):
 * pull-stream state_src with the face overrides, where FOL_new != 0 add the
 * forcing term with u' from FOL_old, take moments with FOL_new, zero u on the
 * masked faces, collide, store into state_dst.  fol_old / fol_new ([3][x][y][z]
 * float) may be NULL (zero).  fol_tiles (uint32 per 256-node tile, written by
 * funlbm_ibm_spread) with ctl limits the FOL reads to the tiles whose entry
 * equals ctl[0] (the last spreading's stamp); NULL reads FOL everywhere.
 * rho_prev / rho_next are used only with WALL_WITH_SPEED faces (the density
 * of the previous step on those faces). */
int funlbm_fused_step(const funlbm_lattice* lat, const void* state_src, void* state_dst,
                      const float* fol_old, const float* fol_new, const uint32_t* fol_tiles,
                      const int32_t* ctl, const float* rho_prev, float* rho_next, void* stream);

/* funlbm_fused_step that also produces the statistics funlbm_flow_stats
 * would read from state_src afterwards (min, mean, max of the reference's f,
 * u and rho after this step's stream and moments, the per-step log of
 * src/funlbm/lbm/base.py:105-144), in the same pass over the state.
 * partials: device double[9 * n_partials], n_partials >= ceil(nodes / 256)
 * + min(ceil(nodes / 256), 256);
 * out: device double[9] as funlbm_flow_stats.  Deterministic: fixed-order
 * reductions.  No reference counterpart (the reference reduces the fields
 * after the step). */
int funlbm_fused_step_stats(const funlbm_lattice* lat, const void* state_src, void* state_dst,
                            const float* fol_old, const float* fol_new, const uint32_t* fol_tiles,
                            const int32_t* ctl, const float* rho_prev, float* rho_next, double* partials,
                            int32_t n_partials, double* out, void* stream);

/* funlbm_fused_step_stats without caller-provided scratch, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The block partials live in a device array inside the library, sized for
 * lattices of at most 2^24 nodes (256^3; 4.5 MiB, held by each GPU the
 * library runs on until the process ends); a larger lattice is
 * FUNLBM_ETOOBIG.  It allocates nothing.  The first call on a device looks up
 * the array's address (hipGetSymbolAddress): make it outside a graph capture.
 * Every call on one GPU shares that GPU's array, so calls that overlap on one
 * GPU (other streams, other threads) race; calls in one stream do not.  The
 * results equal funlbm_fused_step_stats's bit for bit. */
int funlbm_fused_step_stats_noalloc(const funlbm_lattice* lat, const void* state_src, void* state_dst,
                                    const float* fol_old, const float* fol_new, const uint32_t* fol_tiles,
                                    const int32_t* ctl, const float* rho_prev, float* rho_next, double* out,
                                    void* stream);

/* The same pass as funlbm_fused_step without the collision: writes the
 * reference's f, u, rho and p (float) after the step.  Any output may be NULL. */
int funlbm_materialize(const funlbm_lattice* lat, const void* state_src, const float* fol_old,
                       const float* fol_new, const uint32_t* fol_tiles, const int32_t* ctl,
                       const float* rho_prev, float* f, float* u, float* rho, float* p,
                       void* stream);

/* ---- immersed boundary and rigid particles ----
 *
 * Markers of all particles are concatenated particle-major (the reference's
 * loop order, src/funlbm/lbm/lbm3d.py:68,106): lx/lu/lF/lag are float[M][3],
 * lm float[M], marker_particle int32[M], marker_offsets int32[P+1].
 * Particle state is double[P][FUNLBM_PSTRIDE] with the offsets below; `mass`
 * is the volume as in the reference (src/funlbm/particle/base/sphere.py:33-37).
 * ctl is int32[4] device memory: [0] spreading stamp, [1] weight-sum failure
 * flag, [2..3] reserved; zero it once before first use. */
#define FUNLBM_P_CX 0       /* centre [3] */
#define FUNLBM_P_CU 3       /* velocity [3] */
#define FUNLBM_P_CW 6       /* angular velocity [3] */
#define FUNLBM_P_CF 9       /* force [3] */
#define FUNLBM_P_CT 12      /* torque [3] */
#define FUNLBM_P_ANGLE 15   /* accumulated rotation vector [3] */
#define FUNLBM_P_ROT 18     /* rotation matrix of angle, row-major [9] */
#define FUNLBM_P_MASS 27    /* volume */
#define FUNLBM_P_INERTIA 28 /* principal moments [3] */
#define FUNLBM_P_RHO 31     /* particle density */
#define FUNLBM_P_GRAV 32    /* gravity vector [3]; the reference's is (gl, 0, 0) */
#define FUNLBM_PSTRIDE 40

/* Interpolation and direct forcing, replaces LBMD3.flow_to_lagrange
 * (src/funlbm/lbm/lbm3d.py:46-100): per marker, cosine-kernel weights
 * (1 + cos(pi r/2))/4 per axis on the stencil [max(floor(x)-1,0), min(+4,size)),
 * normalised to sum 1; s_u = sum w (ux + uy + uz); s_rho = sum w rho;
 * lu = cu + cw x (lx - cx); lF = 2 s_rho (lu - s_u)/dt.  u and rho come from
 * the fused state (state != NULL: pull + moments with fol_old) or from the
 * explicit fields u/rho.  check_weight_sum != 0 sets ctl[1] for a marker whose
 * weights do not sum to 1 within 1e-3 (the reference checks after step 100). */
int funlbm_ibm_interp(const funlbm_lattice* lat, const void* state, const float* fol_old,
                      const float* rho_face, const float* u, const float* rho, int32_t n_markers,
                      const float* lx, const int32_t* marker_particle, const double* pstate,
                      float dt, int32_t check_weight_sum, float* lu, float* lF, int32_t* ctl,
                      void* stream);

/* funlbm_ibm_interp from the fused state, with the moments of each particle's
 * box computed once per node (the same flow_to_lagrange,
 * src/funlbm/lbm/lbm3d.py:46-100; the results equal funlbm_ibm_interp's bit
 * for bit).  particle_box is int32[n_particles][3] = {offset of the box in
 * box_cache (nodes), R, first chunk}: the box of a particle whose markers lie
 * within R of its centre is the cube of side 2R + 4 from floor(centre) - R - 1,
 * (2R + 4)^3 nodes of 16 bytes from box_cache + 4*offset; R = -1 gives no box
 * (its markers pull per node).  chunk_particle[n_chunks] names the particle of
 * each 256-node chunk, ceil((2R + 4)^3 / 256) consecutive chunks per box from
 * its first chunk.  box_cache holds cache_nodes float4 (16-byte aligned); a box
 * beyond it is not used.  A marker with a stencil node outside its box (no
 * box, a stencil clamped at a face, markers not yet placed) is interpolated
 * per node after the others: miss_list (int32[n_markers + 1]) lists them.
 * The cache and the list are scratch.  n_chunks = 0: every marker per node
 * (box_cache, chunk_particle and miss_list may then be NULL). */
int funlbm_ibm_interp_boxed(const funlbm_lattice* lat, const void* state, const float* fol_old,
                            const float* rho_face, int32_t n_markers, const float* lx,
                            const int32_t* marker_particle, int32_t n_particles,
                            const double* pstate, const int32_t* particle_box, int32_t n_chunks,
                            const int32_t* chunk_particle, float* box_cache, int32_t cache_nodes,
                            int32_t* miss_list, float dt, int32_t check_weight_sum, float* lu,
                            float* lF, int32_t* ctl, void* stream);

/* Wall spring, replaces LBMD3.particle_to_wall (src/funlbm/lbm/lbm3d.py:121-167). */
int funlbm_wall_spring(const funlbm_lattice* lat, int32_t n_particles, const int32_t* marker_offsets,
                       const float* lx, float* lF, float dx, void* stream);

/* Rigid-body update in double, replaces Particle.update_from_lar
 * (src/funlbm/particle/base/base.py:120-151).  rho_fluid is the reference's
 * rouf (1.0); dt <= 0 or rho_fluid <= 0 is FUNLBM_EINVAL. */
int funlbm_particle_forces(int32_t n_particles, const int32_t* marker_offsets, const float* lx,
                           const float* lF, const float* lm, double* pstate, double dt,
                           double rho_fluid, void* stream);

/* Pose update, replaces Particle.update + Coordinate.update/cul_point
 * (src/funlbm/particle/base/base.py:153-156, src/funlbm/particle/coord.py:73-108):
 * if advance, angle += cw; R = rotation of the rotation vector angle;
 * lx = float(R lag) + float(cx). */
int funlbm_particle_pose(int32_t n_particles, const int32_t* marker_offsets, const float* lag,
                         float* lx, double* pstate, int32_t advance, void* stream);

/* Body-frame surface markers of n_shapes ellipsoids, replaces
 * generate_uniform_points_on_ellipsoid / find_intersection
 * (src/funlbm/particle/base/ellipsoid.py:10-95) as called by Sphere._init and
 * Ellipsoid._init (src/funlbm/particle/base/sphere.py:15-53,
 * ellipsoid.py:107-143).  semi: device double[n_shapes][3] semi-axes (a = b = c
 * for a sphere).  markers: device float[n_shapes][cap][3]; the first
 * min(count, cap) rows of each shape are written, in cube order.  counts:
 * device int32[n_shapes], the full marker count of each shape (a count above
 * cap asks for a larger cap), or -1 for a shape with a non-positive or
 * non-finite semi-axis or more than 1024 axis points.  The intersections are
 * computed in closed form in double (the reference runs fsolve from t = 0.5;
 * both give the single in-segment root). */
int funlbm_surface_markers(int32_t n_shapes, const double* semi, double dx, int32_t cap,
                           float* markers, int32_t* counts, void* stream);

/* Spreading, replaces LBMD3.lagrange_to_flow (src/funlbm/lbm/lbm3d.py:102-119):
 * fol := 0 and then, for every marker in order, fol[stencil] = w lF lm
 * overwriting.  The buffer must be zero outside the stencils whose origins
 * (int32[M][3]) are given in clear_origin (NULL: zero everywhere); the origins
 * of this call are written to `origin`.  owner is uint64[nodes] device memory,
 * zeroed once before first use and never again.  fol_tiles (uint32 per
 * 256-node tile, may be NULL) receives this call's stamp in every tile it
 * writes; ctl[0] holds that stamp after the call. */
int funlbm_ibm_spread(const funlbm_lattice* lat, int32_t n_markers, const float* lx, const float* lF,
                      const float* lm, float* fol, const int32_t* clear_origin, int32_t* origin,
                      uint64_t* owner, uint32_t* fol_tiles, int32_t* ctl, void* stream);

/* funlbm_ibm_spread with the claims of each particle whose box (the rows of
 * funlbm_ibm_interp_boxed; a box of at most 16384 nodes) holds its stencils
 * formed in shared memory: a segmented maximum of the marker index per node
 * of the box, then one global claim per node and particle.  The other
 * particles' markers claim per node as in funlbm_ibm_spread.  Same results
 * (the overwrite order of src/funlbm/lbm/lbm3d.py:102-119).
 * marker_offsets[n_particles + 1] bound each particle's markers.
 * claims_per_marker = 0 declares that every particle has such a box
 * (0 <= R <= 10) and skips the per-marker claim kernel; a particle without
 * one then claims nothing. */
int funlbm_ibm_spread_boxed(const funlbm_lattice* lat, int32_t n_markers, const float* lx,
                            const float* lF, const float* lm, float* fol, const int32_t* clear_origin,
                            int32_t* origin, uint64_t* owner, uint32_t* fol_tiles, int32_t* ctl,
                            int32_t n_particles, const int32_t* marker_offsets,
                            const int32_t* marker_particle, const double* pstate,
                            const int32_t* particle_box, int32_t claims_per_marker,
                            void* stream);

/* ---- many lattices per launch ----
 *
 * Independent cases of uneven size share pools: two state pools (the case's
 * [q][n] block at element offset f_offset in each), two FOL pools ([3][n] at
 * 3*node_offset), an owner pool (uint64, n at node_offset) and a tile table
 * (uint32, ceil(n/256) at tile_begin).  A case descriptor is an opaque block
 * of funlbm_case_bytes() bytes built on the host by funlbm_case_pack and
 * copied by the caller into a device array indexed by case slot.  No
 * reference counterpart: the reference solves one lattice per process. */
int funlbm_case_bytes(void);

/* Fills the descriptor of one case.  PERIODICAL and WALL faces only. */
int funlbm_case_pack(const funlbm_lattice* lat, int64_t f_offset, int64_t node_offset,
                     int32_t tile_begin, void* desc_out);

/* One timestep of every case whose steps-left counter is positive: the
 * single-lattice sequence (interp, wall, forces, spreading, fused step, pose)
 * as persistent kernels over device-side work lists, then the per-case
 * counters advance (parity flips, left -1, done +1) and hdr[4] +1.  A case's
 * spreading stamps are its own step counts (done + 1 in the step), so its
 * tile entries and owner words must be zero when it is admitted; after its
 * last step, funlbm_materialize of the case takes ctl[0] = its done count.
 * hdr is int32[9] device memory: [0] active tiles in
 * tile_case/tile_local, [1] marker slots, [2] particle slots, [3] case slots,
 * [4] steps launched, [5], [6], [7] the ends of the D3Q13, D3Q15 and D3Q19 tiles: the
 * tile list holds the tiles of D3Q13 cases first, then D3Q15, D3Q19 and D3Q27
 * ones (one fused kernel per lattice takes its range), [8] entries of
 * chunk_particle.  A slot whose case index is -1 is skipped.  particle_markers
 * holds [begin, end) marker ranges (two int32 per particle).  traj (may be
 * NULL) receives float[4] = (cx, cy, cz, step) per particle and step in a ring
 * of traj_cap[0] entries.  box_cache (may be NULL: every marker pulls per
 * node), particle_box (int32[3] per particle slot), chunk_particle and
 * miss_list (int32, marker slots + 1) are the moment boxes of
 * funlbm_ibm_interp_boxed, over particle slots.  lattice_mask (bits 0-3:
 * D3Q13, D3Q15, D3Q19, D3Q27) selects the fused kernels launched; a lattice
 * whose bit is clear is not stepped.  Bit 4 runs the per-marker spreading
 * claims, needed while some particle has no box of 0 <= R <= 10 (the
 * others claim through their boxes, as in funlbm_ibm_spread_boxed).  The
 * grid sizes fix the persistent launch shapes; lattice_grid (host int32[4],
 * may be NULL: grid_tiles for each) gives the fused kernel of each lattice
 * its own grid, so it does not sweep the other lattices' tiles.  Every grid
 * is a grid-stride loop: any positive size gives the same results. */
int funlbm_batch_step(int32_t storage, const void* cases, const int32_t* hdr, int32_t* case_par,
                      int32_t* case_left, int32_t* case_done, int32_t* case_flag,
                      const int32_t* tile_case, const int32_t* tile_local,
                      const int32_t* marker_case, const int32_t* marker_particle,
                      const int32_t* case_marker_begin, const int32_t* particle_case,
                      const int32_t* particle_markers, void* pool_a, void* pool_b, float* fol_a,
                      float* fol_b, uint64_t* owner, uint32_t* fol_tiles, int32_t* origin_a,
                      int32_t* origin_b, int32_t* origin_ok, const float* lag, float* lx, float* lu,
                      float* lF, const float* lm, double* pstate, float* traj,
                      const int32_t* traj_cap, int32_t grid_tiles, int32_t grid_markers,
                      int32_t grid_particles, int32_t with_ibm, const int32_t* particle_box,
                      const int32_t* chunk_particle, float* box_cache, int32_t cache_nodes,
                      int32_t* miss_list, int32_t lattice_mask, const int32_t* lattice_grid,
                      void* stream);

/* min, mean and max of the reference's f, u and rho after the last fused
 * step, computed from the state without materialising the fields (the
 * statistics of src/funlbm/flow/base.py:183-206 that the run loop logs every
 * step, src/funlbm/lbm/base.py:105-144).  out: double[9] = f min, mean, max,
 * u min, mean, max, rho min, mean, max; the mean of f is the mean of rho
 * over q (a node's populations sum to its density).  partials: device scratch of
 * 9 * n_partials doubles (one entry per block; n_partials >= 1 bounds the
 * grid).  Deterministic: fixed-order reductions. */
int funlbm_flow_stats(const funlbm_lattice* lat, const void* state_src, const float* fol_old,
                      const float* fol_new, const uint32_t* fol_tiles, const int32_t* ctl,
                      const float* rho_prev, double* partials, int32_t n_partials, double* out,
                      void* stream);

/* Global mass and momentum after the last fused step, in double: out[0] =
 * sum of rho, out[1..3] = sum of rho u_d over the nodes (rho and u as the
 * reference's float32 tensors hold them, the forcing correction included).
 * Arguments as funlbm_flow_stats; partials: 4 * n_partials doubles.
 * Deterministic (fixed-order reductions).  No reference counterpart: a
 * conservation check of the fused state. */
int funlbm_flow_conservation(const funlbm_lattice* lat, const void* state_src, const float* fol_old,
                             const float* fol_new, const uint32_t* fol_tiles, const int32_t* ctl,
                             const float* rho_prev, double* partials, int32_t n_partials, double* out,
                             void* stream);

/* Plain device copy of n_bytes (16-byte aligned), used to calibrate the
 * measured copy bandwidth.  No reference counterpart. */
int funlbm_copy_calibrate(const void* src, void* dst, int64_t n_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FUNLBM_H */
