// Immersed-boundary kernels of one lattice and their C-ABI entry points:
// marker interpolation and direct forcing, wall spring, force/torque reduction
// and rigid-body update in double precision, pose update, and spreading with
// the reference's overwrite semantics made deterministic (bodies in
// ibm_body.h).
//
// Markers of all particles are concatenated (particle-major, the reference's
// loop order).  16 lanes serve one marker, four markers per wavefront.
#include "capi_util.h"
#include "ibm_body.h"
#include "lbm_common.h"

namespace funlbm {

template <class L, class S, bool FROM_STATE>
__global__ __launch_bounds__(kIbmBlock) void k_interp(
    KParams p, const typename S::elem* __restrict__ state, const float* __restrict__ fol_old,
    const float* __restrict__ rho_face, const float* __restrict__ ufield,
    const float* __restrict__ rhofield, int nm, const float* __restrict__ lx,
    const int* __restrict__ pid, const double* __restrict__ ps, float dt, int check_sum,
    float* __restrict__ lu, float* __restrict__ lF, int* __restrict__ ctl) {
  const int m = blockIdx.x * kMarkersPerBlock + (int)threadIdx.x / kMarkerLanes;
  const int sub = threadIdx.x & (kMarkerLanes - 1);
  if (m >= nm) return;
  interp_marker<L, S, FROM_STATE>(p, m, sub, state, fol_old, rho_face, ufield, rhofield, lx,
                                  pst(ps, pid[m]), dt, check_sum != 0, lu, lF, &ctl[1]);
}

__global__ __launch_bounds__(kIbmBlock) void k_interp_boxed(
    KParams p, int nm, const float* __restrict__ lx, const int* __restrict__ pid,
    const double* __restrict__ ps, float dt, int check_sum, float* __restrict__ lu,
    float* __restrict__ lF, int* __restrict__ ctl, const int* __restrict__ pbox,
    const float4* __restrict__ cache, int cap, int* __restrict__ miss_list) {
  // consecutive markers (one particle's, sharing box nodes) on one XCD's L2
  const int b = (int)xcd_block(blockIdx.x, gridDim.x);
  const int m = b * kMarkersPerBlock + (int)threadIdx.x / kMarkerLanes;
  const int sub = threadIdx.x & (kMarkerLanes - 1);
  if (m >= nm) return;
  const int pi = pid[m];
  interp_marker_boxed(p, m, sub, lx, pst(ps, pi), dt, check_sum != 0, lu, lF, &ctl[1], pbox + 3 * pi,
                      cache, cap, miss_list);
}

// The markers k_interp_boxed left: per-node interpolation from the state.
template <class L, class S>
__global__ __launch_bounds__(kIbmBlock) void k_interp_missed(
    KParams p, const typename S::elem* __restrict__ state, const float* __restrict__ fol_old,
    const float* __restrict__ rho_face, const int* __restrict__ miss_list,
    const float* __restrict__ lx, const int* __restrict__ pid, const double* __restrict__ ps,
    float dt, int check_sum, float* __restrict__ lu, float* __restrict__ lF, int* __restrict__ ctl) {
  const int n = miss_list[0];
  const int sub = threadIdx.x & (kMarkerLanes - 1);
  for (int e = blockIdx.x * kMarkersPerBlock + (int)threadIdx.x / kMarkerLanes; e < n;
       e += gridDim.x * kMarkersPerBlock) {
    const int m = miss_list[1 + e];
    interp_marker<L, S, true>(p, m, sub, state, fol_old, rho_face, nullptr, nullptr, lx,
                              pst(ps, pid[m]), dt, check_sum != 0, lu, lF, &ctl[1]);
  }
}

// Moment cache of the particles' boxes (box_node), one 256-node chunk per
// block pass.
template <class L, class S>
__global__ __launch_bounds__(kBoxChunk) void k_box_moments(
    KParams p, const typename S::elem* __restrict__ state, const float* __restrict__ fol_old,
    const float* __restrict__ rho_face, const double* __restrict__ ps, const int* __restrict__ pbox,
    const int* __restrict__ chunk_particle, int n_chunks, int np, float4* __restrict__ cache,
    int cap, int* __restrict__ miss_list) {
  if (blockIdx.x == 0 && threadIdx.x == 0) miss_list[0] = 0;  // k_interp_boxed appends
  for (int c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    const int pi = chunk_particle[c];
    if ((unsigned)pi >= (unsigned)np) continue;
    box_node<L, S>(p, state, fol_old, rho_face, pst(ps, pi), pbox + 3 * pi, c, cache, cap);
  }
}

__global__ __launch_bounds__(kIbmBlock) void k_wall(KParams p, const int* __restrict__ offsets,
                                                    const float* __restrict__ lx,
                                                    float* __restrict__ lF, float dx) {
  wall_particle(p, offsets[blockIdx.x], offsets[blockIdx.x + 1], lx, lF, dx);
}

__global__ __launch_bounds__(kIbmBlock) void k_forces(const int* __restrict__ offsets,
                                                      const float* __restrict__ lx,
                                                      const float* __restrict__ lF,
                                                      const float* __restrict__ lm,
                                                      double* __restrict__ ps, double dt,
                                                      double rho_f) {
  forces_particle(offsets[blockIdx.x], offsets[blockIdx.x + 1], lx, lF, lm, pst(ps, blockIdx.x), dt,
                  rho_f);
}

__global__ __launch_bounds__(kIbmBlock) void k_pose(const int* __restrict__ offsets,
                                                    const float* __restrict__ lag,
                                                    float* __restrict__ lx, double* __restrict__ ps,
                                                    int advance) {
  pose_particle(offsets[blockIdx.x], offsets[blockIdx.x + 1], lag, lx, pst(ps, blockIdx.x),
                advance != 0);
}

// pid / pbox (may be NULL): markers of particles that k_spread_claim serves
// are left out.
__global__ __launch_bounds__(kIbmBlock) void k_spread_prep(KParams p, int nm,
                                                           const float* __restrict__ lx,
                                                           const int* __restrict__ clear_origin,
                                                           float* __restrict__ fol,
                                                           unsigned long long* __restrict__ owner,
                                                           const int* __restrict__ ctl,
                                                           const int* __restrict__ pid,
                                                           const int* __restrict__ pbox) {
  __shared__ PrepPeer peers[kIbmBlock / kWave][kMarkersPerWave];
  const int m = blockIdx.x * kMarkersPerBlock + (int)threadIdx.x / kMarkerLanes;
  const int sub = threadIdx.x & (kMarkerLanes - 1);
  const bool active = m < nm && !(pbox && claims_in_box(pbox + 3 * pid[m]));
  // every lane takes part (spread_prep_marker shares the wave's stencils)
  spread_prep_marker(p, active, 0, m, m, sub, lx, clear_origin, fol, owner, (uint32_t)(ctl[0] + 1),
                     peers[threadIdx.x / kWave]);
}

// Claims of the particles with a box of at most kClaimNodes nodes, one block
// per particle (spread_claim_particle).
__global__ __launch_bounds__(kClaimBlock) void k_spread_claim(KParams p, const int* __restrict__ offsets,
                                                            const float* __restrict__ lx,
                                                            const double* __restrict__ ps,
                                                            const int* __restrict__ pbox,
                                                            const int* __restrict__ clear_origin,
                                                            float* __restrict__ fol,
                                                            unsigned long long* __restrict__ owner,
                                                            const int* __restrict__ ctl) {
  __shared__ int own[kClaimNodes];
  const int i = blockIdx.x;
  const int* row = pbox + 3 * i;
  if (!claims_in_box(row)) return;
  spread_claim_particle(p, offsets[i], offsets[i + 1], offsets[i], lx, pst(ps, i), row, clear_origin,
                        fol, owner, (uint32_t)(ctl[0] + 1), own);
}

__global__ __launch_bounds__(kIbmBlock) void k_spread_write(
    KParams p, int nm, const float* __restrict__ lx, const float* __restrict__ lF,
    const float* __restrict__ lm, float* __restrict__ fol,
    const unsigned long long* __restrict__ owner, int* __restrict__ origin,
    uint32_t* __restrict__ tiles, const int* __restrict__ ctl) {
  const int m = blockIdx.x * kMarkersPerBlock + (int)threadIdx.x / kMarkerLanes;
  const int sub = threadIdx.x & (kMarkerLanes - 1);
  if (m >= nm) return;
  spread_write_marker(p, m, m, sub, lx, lF, lm, fol, owner, origin, tiles, (uint32_t)(ctl[0] + 1));
}

__global__ void k_bump(int* ctl) { ctl[0] += 1; }

static int wave_grid(int nm) { return (nm + kMarkersPerBlock - 1) / kMarkersPerBlock; }

}  // namespace funlbm

using namespace funlbm;

extern "C" {

int funlbm_ibm_interp(const funlbm_lattice* lat, const void* state, const float* fol_old,
                      const float* rho_face, const float* u, const float* rho, int32_t n_markers,
                      const float* lx, const int32_t* marker_particle, const double* pstate,
                      float dt, int32_t check_weight_sum, float* lu, float* lF, int32_t* ctl,
                      void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (n_markers < 0) return set_error(FUNLBM_EINVAL, "ibm_interp: negative marker count");
  if (n_markers == 0) return 0;
  const bool from_state = state != nullptr;
  if (from_state && (rc = reject_neq(p, "ibm_interp from a state"))) return rc;
  if (!from_state && (!u || !rho))
    return set_error(FUNLBM_EINVAL, "ibm_interp: give either a state or u and rho");
  if (!lx || !marker_particle || !pstate || !lu || !lF || !ctl)
    return set_error(FUNLBM_EINVAL, "ibm_interp: NULL array");
  if (from_state && p.has_moving && !rho_face)
    return set_error(FUNLBM_EINVAL, "ibm_interp: moving wall needs the face density");
  if (!(dt > 0.0f)) return set_error(FUNLBM_EINVAL, "ibm_interp: dt must be positive");
  hipStream_t s = (hipStream_t)stream;
  return dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    return dispatch_storage(from_state ? lat->storage : FUNLBM_STORE_F32, [&](auto st) {
      using S = decltype(st);
      using E = typename S::elem;
      if (from_state)
        hipLaunchKernelGGL((k_interp<L, S, true>), dim3(wave_grid(n_markers)), dim3(kIbmBlock), 0, s, p,
                           (const E*)state, fol_old, rho_face, u, rho, n_markers, lx, marker_particle,
                           pstate, dt, check_weight_sum, lu, lF, ctl);
      else
        hipLaunchKernelGGL((k_interp<L, S, false>), dim3(wave_grid(n_markers)), dim3(kIbmBlock), 0, s,
                           p, (const E*)state, fol_old, rho_face, u, rho, n_markers, lx,
                           marker_particle, pstate, dt, check_weight_sum, lu, lF, ctl);
      return check_launch("ibm_interp");
    });
  });
}

int funlbm_ibm_interp_boxed(const funlbm_lattice* lat, const void* state, const float* fol_old,
                            const float* rho_face, int32_t n_markers, const float* lx,
                            const int32_t* marker_particle, int32_t n_particles,
                            const double* pstate, const int32_t* particle_box, int32_t n_chunks,
                            const int32_t* chunk_particle, float* box_cache, int32_t cache_nodes,
                            int32_t* miss_list, float dt, int32_t check_weight_sum, float* lu,
                            float* lF, int32_t* ctl, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (n_markers < 0 || n_particles < 0 || n_chunks < 0 || cache_nodes < 0)
    return set_error(FUNLBM_EINVAL, "ibm_interp_boxed: negative count");
  if (n_markers == 0) return 0;
  if ((rc = reject_neq(p, "ibm_interp from a state"))) return rc;
  if (!state || !lx || !marker_particle || !pstate || !particle_box || !lu || !lF || !ctl ||
      (n_chunks > 0 && (!chunk_particle || !box_cache || !miss_list)))
    return set_error(FUNLBM_EINVAL, "ibm_interp_boxed: NULL array");
  if (box_cache && ((uintptr_t)box_cache & 15))
    return set_error(FUNLBM_EINVAL, "ibm_interp_boxed: box cache must be 16-byte aligned");
  if (p.has_moving && !rho_face)
    return set_error(FUNLBM_EINVAL, "ibm_interp: moving wall needs the face density");
  if (!(dt > 0.0f)) return set_error(FUNLBM_EINVAL, "ibm_interp: dt must be positive");
  hipStream_t s = (hipStream_t)stream;
  return dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    return dispatch_storage(lat->storage, [&](auto st) {
      using S = decltype(st);
      using E = typename S::elem;
      if (n_chunks == 0) {  // no box: every marker per node
        hipLaunchKernelGGL((k_interp<L, S, true>), dim3(wave_grid(n_markers)), dim3(kIbmBlock), 0, s,
                           p, (const E*)state, fol_old, rho_face, nullptr, nullptr, n_markers, lx,
                           marker_particle, pstate, dt, check_weight_sum, lu, lF, ctl);
        return check_launch("ibm_interp");
      }
      hipLaunchKernelGGL((k_box_moments<L, S>), dim3(n_chunks < 65536 ? n_chunks : 65536),
                         dim3(kBoxChunk), 0, s, p, (const E*)state, fol_old, rho_face, pstate,
                         particle_box, chunk_particle, n_chunks, n_particles, (float4*)box_cache,
                         cache_nodes, miss_list);
      if ((rc = check_launch("ibm_box_moments"))) return rc;
      hipLaunchKernelGGL(k_interp_boxed, dim3(wave_grid(n_markers)), dim3(kIbmBlock), 0, s, p,
                         n_markers, lx, marker_particle, pstate, dt, check_weight_sum, lu, lF, ctl,
                         particle_box, (const float4*)box_cache, cache_nodes, miss_list);
      if ((rc = check_launch("ibm_interp_boxed"))) return rc;
      // the markers left over (their count is on the device): a small fixed grid
      hipLaunchKernelGGL((k_interp_missed<L, S>), dim3(64), dim3(kIbmBlock), 0, s, p, (const E*)state,
                         fol_old, rho_face, (const int*)miss_list, lx, marker_particle, pstate, dt,
                         check_weight_sum, lu, lF, ctl);
      return check_launch("ibm_interp_missed");
    });
  });
}

int funlbm_wall_spring(const funlbm_lattice* lat, int32_t n_particles, const int32_t* marker_offsets,
                       const float* lx, float* lF, float dx, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (n_particles <= 0) return n_particles == 0 ? 0 : set_error(FUNLBM_EINVAL, "negative count");
  if (!marker_offsets || !lx || !lF) return set_error(FUNLBM_EINVAL, "wall_spring: NULL array");
  hipLaunchKernelGGL(k_wall, dim3(n_particles), dim3(kIbmBlock), 0, (hipStream_t)stream, p,
                     marker_offsets, lx, lF, dx);
  return check_launch("wall_spring");
}

int funlbm_particle_forces(int32_t n_particles, const int32_t* marker_offsets, const float* lx,
                           const float* lF, const float* lm, double* pstate, double dt, double rho_fluid,
                           void* stream) {
  if (n_particles <= 0) return n_particles == 0 ? 0 : set_error(FUNLBM_EINVAL, "negative count");
  if (!(dt > 0.0)) return set_error(FUNLBM_EINVAL, "Time step must be positive");
  if (!(rho_fluid > 0.0)) return set_error(FUNLBM_EINVAL, "Fluid density must be positive");
  if (!marker_offsets || !lx || !lF || !lm || !pstate)
    return set_error(FUNLBM_EINVAL, "particle_forces: NULL array");
  hipLaunchKernelGGL(k_forces, dim3(n_particles), dim3(kIbmBlock), 0, (hipStream_t)stream,
                     marker_offsets, lx, lF, lm, pstate, dt, rho_fluid);
  return check_launch("particle_forces");
}

int funlbm_particle_pose(int32_t n_particles, const int32_t* marker_offsets, const float* lag,
                         float* lx, double* pstate, int32_t advance, void* stream) {
  if (n_particles <= 0) return n_particles == 0 ? 0 : set_error(FUNLBM_EINVAL, "negative count");
  if (!marker_offsets || !lag || !lx || !pstate)
    return set_error(FUNLBM_EINVAL, "particle_pose: NULL array");
  hipLaunchKernelGGL(k_pose, dim3(n_particles), dim3(kIbmBlock), 0, (hipStream_t)stream,
                     marker_offsets, lag, lx, pstate, advance);
  return check_launch("particle_pose");
}

int funlbm_ibm_spread(const funlbm_lattice* lat, int32_t n_markers, const float* lx, const float* lF,
                      const float* lm, float* fol, const int32_t* clear_origin, int32_t* origin,
                      uint64_t* owner, uint32_t* fol_tiles, int32_t* ctl, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (n_markers < 0) return set_error(FUNLBM_EINVAL, "ibm_spread: negative marker count");
  if (!fol || !owner || !ctl) return set_error(FUNLBM_EINVAL, "ibm_spread: NULL array");
  if (n_markers > 0 && (!lx || !lF || !lm)) return set_error(FUNLBM_EINVAL, "ibm_spread: NULL array");
  hipStream_t s = (hipStream_t)stream;
  if (n_markers > 0) {
    hipLaunchKernelGGL(k_spread_prep, dim3(wave_grid(n_markers)), dim3(kIbmBlock), 0, s, p, n_markers,
                       lx, clear_origin, fol, (unsigned long long*)owner, ctl, nullptr, nullptr);
    if ((rc = check_launch("ibm_spread"))) return rc;
    hipLaunchKernelGGL(k_spread_write, dim3(wave_grid(n_markers)), dim3(kIbmBlock), 0, s, p, n_markers,
                       lx, lF, lm, fol, (const unsigned long long*)owner, origin, fol_tiles, ctl);
    if ((rc = check_launch("ibm_spread"))) return rc;
  }
  hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, s, ctl);
  return check_launch("ibm_spread");
}

int funlbm_ibm_spread_boxed(const funlbm_lattice* lat, int32_t n_markers, const float* lx,
                            const float* lF, const float* lm, float* fol, const int32_t* clear_origin,
                            int32_t* origin, uint64_t* owner, uint32_t* fol_tiles, int32_t* ctl,
                            int32_t n_particles, const int32_t* marker_offsets,
                            const int32_t* marker_particle, const double* pstate,
                            const int32_t* particle_box, int32_t claims_per_marker,
                            void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (n_markers < 0 || n_particles < 0)
    return set_error(FUNLBM_EINVAL, "ibm_spread_boxed: negative count");
  if (!fol || !owner || !ctl) return set_error(FUNLBM_EINVAL, "ibm_spread_boxed: NULL array");
  if (n_markers > 0 && (!lx || !lF || !lm || !marker_offsets || !marker_particle || !pstate ||
                        !particle_box))
    return set_error(FUNLBM_EINVAL, "ibm_spread_boxed: NULL array");
  hipStream_t s = (hipStream_t)stream;
  if (n_markers > 0) {
    if (n_particles > 0) {
      hipLaunchKernelGGL(k_spread_claim, dim3(n_particles), dim3(kClaimBlock), 0, s, p, marker_offsets, lx,
                         pstate, particle_box, clear_origin, fol, (unsigned long long*)owner, ctl);
      if ((rc = check_launch("ibm_spread_boxed"))) return rc;
    }
    if (claims_per_marker) {
      hipLaunchKernelGGL(k_spread_prep, dim3(wave_grid(n_markers)), dim3(kIbmBlock), 0, s, p, n_markers,
                         lx, clear_origin, fol, (unsigned long long*)owner, ctl, marker_particle,
                         particle_box);
      if ((rc = check_launch("ibm_spread_boxed"))) return rc;
    }
    hipLaunchKernelGGL(k_spread_write, dim3(wave_grid(n_markers)), dim3(kIbmBlock), 0, s, p, n_markers,
                       lx, lF, lm, fol, (const unsigned long long*)owner, origin, fol_tiles, ctl);
    if ((rc = check_launch("ibm_spread_boxed"))) return rc;
  }
  hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, s, ctl);
  return check_launch("ibm_spread_boxed");
}

}  // extern "C"
