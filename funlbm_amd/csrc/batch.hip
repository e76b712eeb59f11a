// Many independent lattices of uneven size stepped together.
//
// Every case owns contiguous blocks of shared pools: [q][n] populations in
// each of the two state pools, [3][n] forces in each of the two FOL pools,
// n owner words, and ceil(n/256) entries of the tile table.  A case
// descriptor (CaseDesc, built on the host by funlbm_case_pack) carries its
// lattice parameters and offsets.  Per-case device counters hold the step
// parity (which pool is the source), the steps left and the steps done; a
// case whose steps-left reach zero stops moving while the others go on.
//
// The kernels are persistent grid-stride loops over device-side work lists
// (active tiles, markers, particles, cases) whose lengths live in a device
// header, so one captured HIP graph of a step sequence stays valid while the
// host admits and retires cases between replays.  Per step:
//   interp -> wall -> forces -> spread_prep -> spread_write -> fused -> pose -> advance
#include "capi_util.h"
#include "fused_node.h"
#include "ibm_body.h"
#include "lbm_common.h"

namespace funlbm {

// Device header (int32): counts of the work lists and of the launched steps.
enum : int {
  H_TILES = 0,      // active tiles in tile_case / tile_local
  H_MARKERS = 1,    // markers in the marker arrays
  H_PARTICLES = 2,  // particles in the particle arrays
  H_CASES = 3,      // cases in the case arrays
  H_STAMP = 4,      // steps launched (each case's spreadings carry its own stamps)
  H_QEND = 5,       // [5], [6], [7]: end of the D3Q13, D3Q15, D3Q19 tiles (the list is
                    // ordered by lattice; D3Q27 tiles run to hdr[H_TILES])
  H_CHUNKS = 8,     // 256-node chunks of the particles' moment boxes
  H_WORDS = 9
};

struct CaseDesc {
  KParams p;
  int32_t q;
  int32_t tile_begin;  // first entry of the case in the tile table
  int32_t n_tiles;
  int32_t pad0;
  int64_t f_off;     // element offset of the [q][n] block in each state pool
  int64_t node_off;  // node offset: FOL blocks at 3*node_off, owner at node_off
};

struct BatchDyn {
  int32_t* par;   // [C] 0: state in pool A / FOL_old in FOL A; 1: the other way
  int32_t* left;  // [C] steps still to take
  int32_t* done;  // [C] steps taken
  int32_t* flag;  // [C] weight-sum failure flag
};

// The spreading stamp of a case's coming step: its own step count plus one.
// A case's tile entries and owner words carry its own stamps, so they do not
// depend on how far other cases have stepped (the host zeroes both ranges
// when it admits a case into them).
__device__ inline uint32_t case_stamp(const BatchDyn& dyn, int c) { return (uint32_t)dyn.done[c] + 1u; }

template <class Fn>
__device__ inline void with_lattice(int q, Fn&& fn) {
  switch (q) {
    case 13: fn(D3Q13{}); break;
    case 15: fn(D3Q15{}); break;
    case 19: fn(D3Q19{}); break;
    default: fn(D3Q27{}); break;
  }
}

// ---------------------------------------------------------------------------
// Fused step over the active tiles (256 nodes of one case each).
// ---------------------------------------------------------------------------
// One kernel per lattice (L), each taking the tiles of its own cases: a
// single kernel over all four lattices carries the registers of the largest.
template <class S, class L>
__global__ __launch_bounds__(kBlock) void k_batch_fused(
    const CaseDesc* __restrict__ cases, const int* __restrict__ tile_case,
    const int* __restrict__ tile_local, const int* __restrict__ hdr, const int* __restrict__ case_left,
    const int* __restrict__ case_par, const int* __restrict__ case_done,
    typename S::elem* __restrict__ pool_a, typename S::elem* __restrict__ pool_b,
    const float* __restrict__ fol_a, const float* __restrict__ fol_b, const uint32_t* __restrict__ fol_tiles) {
  using E = typename S::elem;
  constexpr int li = L::Q == 13 ? 0 : L::Q == 15 ? 1 : L::Q == 19 ? 2 : 3;
  const int t_begin = li == 0 ? 0 : hdr[H_QEND + li - 1];
  const int t_end = li == 3 ? hdr[H_TILES] : hdr[H_QEND + li];
  // consecutive tiles (one case's, whose shifted rows share cache lines) on
  // one XCD's L2: the hardware deals blocks round-robin over the XCDs
  const int g = (int)xcd_block(blockIdx.x, gridDim.x);
  for (int t = t_begin + g; t < t_end; t += gridDim.x) {
    // the tile's case and place, then everything that depends on the case
    // only, loaded together before any test: one round trip each (scalar
    // loads; this kernel writes none of these arrays), not a chain of
    // tests and loads.  A free tile entry (-1) reads case slot 0 and skips.
    const int c_raw = tile_case[t];
    const int lt = tile_local[t];
    const int c = c_raw < 0 ? 0 : c_raw;
    const CaseDesc& cd = cases[c];
    const int left = case_left[c];
    const int par = case_par[c];
    const uint32_t done = (uint32_t)case_done[c];
    const int q = cd.q;
    const int64_t n = cd.p.n;
    const int64_t f_off = cd.f_off;
    const int tile_begin = cd.tile_begin;
    // keeps the compiler from sinking each load behind the test before it
    // (pinning the pull table as well costs a wave of occupancy: slower)
    __asm__ volatile("" ::"s"(left), "s"(par), "s"(done), "s"(q), "s"(n), "s"(f_off), "s"(tile_begin));
    if (c_raw < 0 || left <= 0 || q != L::Q) continue;
    const int64_t node = (int64_t)lt * kBlock + threadIdx.x;
    if (node >= n) continue;
    const uint32_t stamp = done + 1u;  // this step's spreading (case_stamp)
    E* pa = pool_a + f_off;
    E* pb = pool_b + f_off;
    const E* src = par ? pb : pa;
    E* dst = par ? pa : pb;
    const float* fa = fol_a + 3 * cd.node_off;
    const float* fb = fol_b + 3 * cd.node_off;
    const float* fo = par ? fb : fa;
    const float* fn = par ? fa : fb;
    const uint32_t entry = fol_tiles[tile_begin + lt];
    const bool new_here = entry == stamp;
    const bool old_here = entry == stamp || entry == stamp - 1u;
    FusedIO<S> io{src, dst, fo, fn, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // the case's dither seed advances with its own step count, as a
    // single-lattice solve's does with its steps
    const uint32_t mix = mix32(step_seed(cd.p.seed, done));
    fused_step_node<L, S, true>(cd.p, node, io, new_here, old_here, mix);
  }
}

// ---------------------------------------------------------------------------
// IBM over all markers / particles of stepping cases.
// ---------------------------------------------------------------------------
// Per-node interpolation from the state: of every marker (list == NULL) or
// of the markers in list ([0] = count, then marker indices).
template <class S>
__global__ __launch_bounds__(kIbmBlock) void k_batch_interp_pull(
    const CaseDesc* __restrict__ cases, const int* __restrict__ hdr, BatchDyn dyn,
    const int* __restrict__ marker_case, const int* __restrict__ marker_particle,
    const typename S::elem* __restrict__ pool_a, const typename S::elem* __restrict__ pool_b,
    const float* __restrict__ fol_a, const float* __restrict__ fol_b,
    const float* __restrict__ lx, const double* __restrict__ ps, float* __restrict__ lu,
    float* __restrict__ lF, const int* __restrict__ list) {
  const int n = list ? list[0] : hdr[H_MARKERS];
  const int sub = threadIdx.x & (kMarkerLanes - 1);
  const int stride = gridDim.x * kMarkersPerBlock;
  for (int e = blockIdx.x * kMarkersPerBlock + (int)threadIdx.x / kMarkerLanes; e < n; e += stride) {
    const int m = list ? list[1 + e] : e;
    const int c = marker_case[m];
    if (c < 0 || dyn.left[c] <= 0) continue;
    const CaseDesc& cd = cases[c];
    const int par = dyn.par[c];
    const typename S::elem* st = (par ? pool_b : pool_a) + cd.f_off;
    const float* fo = (par ? fol_b : fol_a) + 3 * cd.node_off;
    const bool check = dyn.done[c] + 1 > 100;
    with_lattice(cd.q, [&](auto l) {
      using L = decltype(l);
      interp_marker<L, S, true>(cd.p, m, sub, st, fo, nullptr, nullptr, nullptr, lx,
                                pst(ps, marker_particle[m]), 1.0f, check, lu, lF, &dyn.flag[c]);
    });
  }
}

// Interpolation from the particles' moment boxes; markers with a node outside
// their box go to miss_list for k_batch_interp_pull.
__global__ __launch_bounds__(kIbmBlock) void k_batch_interp_boxed(
    const CaseDesc* __restrict__ cases, const int* __restrict__ hdr, BatchDyn dyn,
    const int* __restrict__ marker_case, const int* __restrict__ marker_particle,
    const float* __restrict__ lx, const double* __restrict__ ps, float* __restrict__ lu,
    float* __restrict__ lF, const int* __restrict__ pbox, const float4* __restrict__ cache, int cap,
    int* __restrict__ miss_list) {
  const int nm = hdr[H_MARKERS];
  const int sub = threadIdx.x & (kMarkerLanes - 1);
  const int stride = gridDim.x * kMarkersPerBlock;
  // consecutive markers (one particle's, sharing box nodes) on one XCD's L2
  const int b = (int)xcd_block(blockIdx.x, gridDim.x);
  for (int m = b * kMarkersPerBlock + (int)threadIdx.x / kMarkerLanes; m < nm; m += stride) {
    const int c = marker_case[m];
    if (c < 0 || dyn.left[c] <= 0) continue;
    const int pi = marker_particle[m];
    interp_marker_boxed(cases[c].p, m, sub, lx, pst(ps, pi), 1.0f, dyn.done[c] + 1 > 100, lu, lF,
                        &dyn.flag[c], pbox + 3 * pi, cache, cap, miss_list);
  }
}

// Moment boxes of the active cases' particles (box_node), before
// k_batch_interp_boxed reads them.  chunk_particle names each chunk's particle.
template <class S>
__global__ __launch_bounds__(kBoxChunk) void k_batch_box_moments(
    const CaseDesc* __restrict__ cases, const int* __restrict__ hdr, BatchDyn dyn,
    const int* __restrict__ particle_case, const typename S::elem* __restrict__ pool_a,
    const typename S::elem* __restrict__ pool_b, const float* __restrict__ fol_a,
    const float* __restrict__ fol_b, const double* __restrict__ ps, const int* __restrict__ pbox,
    const int* __restrict__ chunk_particle, float4* __restrict__ cache, int cap,
    int* __restrict__ miss_list) {
  if (blockIdx.x == 0 && threadIdx.x == 0) miss_list[0] = 0;  // k_batch_interp_boxed appends
  const int nch = hdr[H_CHUNKS];
  const int np = hdr[H_PARTICLES];
  for (int ch = blockIdx.x; ch < nch; ch += gridDim.x) {
    const int pi = chunk_particle[ch];
    if ((unsigned)pi >= (unsigned)np) continue;
    const int c = particle_case[pi];
    if (c < 0 || dyn.left[c] <= 0) continue;
    const CaseDesc& cd = cases[c];
    const int par = dyn.par[c];
    const typename S::elem* st = (par ? pool_b : pool_a) + cd.f_off;
    const float* fo = (par ? fol_b : fol_a) + 3 * cd.node_off;
    with_lattice(cd.q, [&](auto l) {
      using L = decltype(l);
      box_node<L, S>(cd.p, st, fo, nullptr, pst(ps, pi), pbox + 3 * pi, ch, cache, cap);
    });
  }
}

// particle_markers holds [begin, end) marker ranges, two int32 per particle.
// Wall spring, then force/torque reduction and rigid-body update, of each
// active particle (one block per particle; the spring's lF rows are read by
// the reduction after the block barrier).
__global__ __launch_bounds__(kIbmBlock) void k_batch_wall_forces(
    const CaseDesc* __restrict__ cases, const int* __restrict__ hdr, BatchDyn dyn,
    const int* __restrict__ particle_case, const int* __restrict__ pm, const float* __restrict__ lx,
    float* __restrict__ lF, const float* __restrict__ lm, double* __restrict__ ps, float dx) {
  const int np = hdr[H_PARTICLES];
  for (int i = blockIdx.x; i < np; i += gridDim.x) {
    const int c = particle_case[i];
    if (c < 0 || dyn.left[c] <= 0) continue;
    wall_particle(cases[c].p, pm[2 * i], pm[2 * i + 1], lx, lF, dx);
    __syncthreads();
    forces_particle(pm[2 * i], pm[2 * i + 1], lx, lF, lm, pst(ps, i), 1.0, 1.0);
    __syncthreads();  // shared memory is reused by the next particle
  }
}

__global__ __launch_bounds__(kIbmBlock) void k_batch_spread_prep(
    const CaseDesc* __restrict__ cases, const int* __restrict__ hdr, BatchDyn dyn,
    const int* __restrict__ marker_case, const int* __restrict__ case_marker_begin,
    const float* __restrict__ lx, const int* __restrict__ origin_a, const int* __restrict__ origin_b,
    const int* __restrict__ origin_ok, float* __restrict__ fol_a, float* __restrict__ fol_b,
    unsigned long long* __restrict__ owner, const int* __restrict__ marker_particle,
    const int* __restrict__ pbox) {
  __shared__ PrepPeer peers[kIbmBlock / kWave][kMarkersPerWave];
  const int nm = hdr[H_MARKERS];
  const int sub = threadIdx.x & (kMarkerLanes - 1);
  const int wv = threadIdx.x / kWave;
  const int g = (threadIdx.x & (kWave - 1)) / kMarkerLanes;
  const int stride = gridDim.x * kMarkersPerBlock;
  // the loop bound is uniform over the wave: its four markers take part in
  // every pass together (spread_prep_marker shares their stencils)
  for (int mb = blockIdx.x * kMarkersPerBlock + wv * kMarkersPerWave; mb < nm; mb += stride) {
    const int m = mb + g;
    int c = -1;
    bool act = m < nm;
    if (act) {
      c = marker_case[m];
      // the markers of a particle k_batch_spread_claim serves are left out
      act = c >= 0 && dyn.left[c] > 0 && !(pbox && claims_in_box(pbox + 3 * marker_particle[m]));
    }
    const CaseDesc& cd = cases[act ? c : 0];
    float* fn = nullptr;
    const int* org = nullptr;
    int key = 0;
    uint32_t stamp = 0;
    if (act) {
      stamp = case_stamp(dyn, c);
      const int par = dyn.par[c];
      // FOL_new of this step is buffer 1 - par; its origins were recorded the
      // last time the case wrote that buffer
      fn = (par ? fol_a : fol_b) + 3 * cd.node_off;
      const bool ok = (origin_ok[c] >> (1 - par)) & 1;
      org = ok ? (par ? origin_a : origin_b) : nullptr;
      key = m - case_marker_begin[c];
    }
    spread_prep_marker(cd.p, act, c, m, key, sub, lx, org, fn, owner + cd.node_off, stamp, peers[wv]);
  }
}

// Claims of the active particles with a box of at most kClaimNodes nodes,
// one block per particle (spread_claim_particle).
__global__ __launch_bounds__(kClaimBlock) void k_batch_spread_claim(
    const CaseDesc* __restrict__ cases, const int* __restrict__ hdr, BatchDyn dyn,
    const int* __restrict__ particle_case, const int* __restrict__ pm,
    const int* __restrict__ case_marker_begin, const float* __restrict__ lx,
    const double* __restrict__ ps, const int* __restrict__ pbox, const int* __restrict__ origin_a,
    const int* __restrict__ origin_b, const int* __restrict__ origin_ok, float* __restrict__ fol_a,
    float* __restrict__ fol_b, unsigned long long* __restrict__ owner) {
  __shared__ int own[kClaimNodes];
  const int np = hdr[H_PARTICLES];
  for (int i = blockIdx.x; i < np; i += gridDim.x) {
    const int c = particle_case[i];
    if (c < 0 || dyn.left[c] <= 0) continue;
    const uint32_t stamp = case_stamp(dyn, c);
    const int* row = pbox + 3 * i;
    if (!claims_in_box(row)) continue;
    const CaseDesc& cd = cases[c];
    const int par = dyn.par[c];
    float* fn = (par ? fol_a : fol_b) + 3 * cd.node_off;
    const bool ok = (origin_ok[c] >> (1 - par)) & 1;
    const int* org = ok ? (par ? origin_a : origin_b) : nullptr;
    const int b = pm[2 * i], e = pm[2 * i + 1];
    spread_claim_particle(cd.p, b, e, b - case_marker_begin[c], lx, pst(ps, i), row, org, fn,
                          owner + cd.node_off, stamp, own);
    __syncthreads();  // own is reused by the next particle
  }
}

__global__ __launch_bounds__(kIbmBlock) void k_batch_spread_write(
    const CaseDesc* __restrict__ cases, const int* __restrict__ hdr, BatchDyn dyn,
    const int* __restrict__ marker_case, const int* __restrict__ case_marker_begin,
    const float* __restrict__ lx, const float* __restrict__ lF, const float* __restrict__ lm,
    int* __restrict__ origin_a, int* __restrict__ origin_b, int* __restrict__ origin_ok,
    float* __restrict__ fol_a, float* __restrict__ fol_b,
    const unsigned long long* __restrict__ owner, uint32_t* __restrict__ fol_tiles) {
  const int nm = hdr[H_MARKERS];
  const int sub = threadIdx.x & (kMarkerLanes - 1);
  const int stride = gridDim.x * kMarkersPerBlock;
  for (int m = blockIdx.x * kMarkersPerBlock + (int)threadIdx.x / kMarkerLanes; m < nm; m += stride) {
    const int c = marker_case[m];
    if (c < 0 || dyn.left[c] <= 0) continue;
    const CaseDesc& cd = cases[c];
    const int par = dyn.par[c];
    float* fn = (par ? fol_a : fol_b) + 3 * cd.node_off;
    int* org = par ? origin_a : origin_b;
    spread_write_marker(cd.p, m, m - case_marker_begin[c], sub, lx, lF, lm, fn, owner + cd.node_off,
                        org, fol_tiles + cd.tile_begin, case_stamp(dyn, c));
    if (sub == 0 && m == case_marker_begin[c]) atomicOr(&origin_ok[c], 1 << (1 - par));
  }
}

__global__ __launch_bounds__(kIbmBlock) void k_batch_pose(const int* __restrict__ hdr, BatchDyn dyn,
                                                          const int* __restrict__ particle_case,
                                                          const int* __restrict__ pm,
                                                          const float* __restrict__ lag,
                                                          float* __restrict__ lx,
                                                          double* __restrict__ ps,
                                                          float* __restrict__ traj,
                                                          const int* __restrict__ traj_cap) {
  const int np = hdr[H_PARTICLES];
  for (int i = blockIdx.x; i < np; i += gridDim.x) {
    const int c = particle_case[i];
    if (c < 0 || dyn.left[c] <= 0) continue;
    double* P = pst(ps, i);
    pose_particle(pm[2 * i], pm[2 * i + 1], lag, lx, P, true);
    if (traj && threadIdx.x == 0) {
      // centroid after this step, ring of traj_cap[0] entries per particle
      const int cap = traj_cap[0];
      const int slot = dyn.done[c] % cap;
      float* t = traj + ((int64_t)i * cap + slot) * 4;
      t[0] = (float)P[FUNLBM_P_CX];
      t[1] = (float)P[FUNLBM_P_CX + 1];
      t[2] = (float)P[FUNLBM_P_CX + 2];
      t[3] = (float)(dyn.done[c] + 1);
    }
  }
}

// Case bookkeeping after a step: parity flips, counters move (a case's step
// count is also the stamp of its last spreading), the launch count advances.
__global__ void k_batch_advance(int* __restrict__ hdr, BatchDyn dyn) {
  const int nc = hdr[H_CASES];
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    if (dyn.left[c] > 0) {
      dyn.par[c] ^= 1;
      dyn.left[c] -= 1;
      dyn.done[c] += 1;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) hdr[H_STAMP] += 1;
}

}  // namespace funlbm

using namespace funlbm;

extern "C" {

int funlbm_case_bytes(void) { return (int)sizeof(CaseDesc); }

int funlbm_case_pack(const funlbm_lattice* lat, int64_t f_offset, int64_t node_offset,
                     int32_t tile_begin, void* desc_out) {
  if (!desc_out) return set_error(FUNLBM_EINVAL, "case_pack: NULL output");
  CaseDesc d;
  memset(&d, 0, sizeof(d));
  int rc = make_params(lat, &d.p);
  if (rc) return rc;
  if ((rc = reject_neq(d.p, "case_pack"))) return rc;
  if (d.p.has_moving)
    return set_error(FUNLBM_EUNSUPPORTED_BC, "batched cases take PERIODICAL and WALL faces only");
  if (f_offset < 0 || node_offset < 0 || tile_begin < 0)
    return set_error(FUNLBM_EINVAL, "case_pack: negative offset");
  d.q = lat->q;
  d.tile_begin = tile_begin;
  d.n_tiles = (int32_t)((d.p.n + kBlock - 1) / kBlock);
  d.f_off = f_offset;
  d.node_off = node_offset;
  memcpy(desc_out, &d, sizeof(d));
  return 0;
}

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
                      void* stream) {
  if (!cases || !hdr || !case_par || !case_left || !case_done || !case_flag || !tile_case ||
      !tile_local || !pool_a || !pool_b || !fol_a || !fol_b || !fol_tiles)
    return set_error(FUNLBM_EINVAL, "batch_step: NULL array");
  if (grid_tiles <= 0 || grid_markers <= 0 || grid_particles <= 0)
    return set_error(FUNLBM_EINVAL, "batch_step: grid sizes must be positive");
  if (lattice_grid)
    for (int i = 0; i < 4; ++i)
      if (((lattice_mask >> i) & 1) && lattice_grid[i] <= 0)
        return set_error(FUNLBM_EINVAL, "batch_step: lattice grid %d must be positive", i);
  if (with_ibm && (!marker_case || !marker_particle || !case_marker_begin || !particle_case ||
                   !particle_markers || !owner || !origin_a || !origin_b || !origin_ok || !lag ||
                   !lx || !lu || !lF || !lm || !pstate))
    return set_error(FUNLBM_EINVAL, "batch_step: NULL particle array");
  if (with_ibm && box_cache && (!particle_box || !chunk_particle || !miss_list || cache_nodes < 0))
    return set_error(FUNLBM_EINVAL, "batch_step: box cache without its tables");
  if (box_cache && ((uintptr_t)box_cache & 15))
    return set_error(FUNLBM_EINVAL, "batch_step: box cache must be 16-byte aligned");
  const float4* cache = (with_ibm && box_cache) ? (const float4*)box_cache : nullptr;
  const CaseDesc* cd = (const CaseDesc*)cases;
  int32_t* h = (int32_t*)hdr;
  BatchDyn dyn{case_par, case_left, case_done, case_flag};
  hipStream_t s = (hipStream_t)stream;
  int rc = 0;
  return dispatch_storage(storage, [&](auto st) {
    using S = decltype(st);
    using E = typename S::elem;
    if (with_ibm) {
      if (cache) {
        hipLaunchKernelGGL((k_batch_box_moments<S>), dim3(grid_markers), dim3(kBoxChunk), 0, s, cd,
                           hdr, dyn, particle_case, (const E*)pool_a, (const E*)pool_b, fol_a, fol_b,
                           pstate, particle_box, chunk_particle, (float4*)box_cache, cache_nodes,
                           miss_list);
        if ((rc = check_launch("batch_box_moments"))) return rc;
        hipLaunchKernelGGL(k_batch_interp_boxed, dim3(grid_markers), dim3(kIbmBlock), 0, s, cd, hdr,
                           dyn, marker_case, marker_particle, lx, pstate, lu, lF, particle_box, cache,
                           cache_nodes, miss_list);
        if ((rc = check_launch("batch_interp_boxed"))) return rc;
      }
      // the markers without a box (all of them without a cache), per node;
      // the count of the list is on the device: a small fixed grid
      hipLaunchKernelGGL((k_batch_interp_pull<S>), dim3(cache ? 64 : grid_markers), dim3(kIbmBlock), 0,
                         s, cd, hdr, dyn, marker_case, marker_particle, (const E*)pool_a,
                         (const E*)pool_b, fol_a, fol_b, lx, pstate, lu, lF,
                         cache ? (const int*)miss_list : nullptr);
      if ((rc = check_launch("batch_interp_pull"))) return rc;
      hipLaunchKernelGGL(k_batch_wall_forces, dim3(grid_particles), dim3(kIbmBlock), 0, s, cd, hdr, dyn,
                         particle_case, particle_markers, lx, lF, lm, pstate, 1.0f);
      if ((rc = check_launch("batch_wall_forces"))) return rc;
      if (particle_box) {
        hipLaunchKernelGGL(k_batch_spread_claim, dim3(grid_particles), dim3(kClaimBlock), 0, s, cd, hdr,
                           dyn, particle_case, particle_markers, case_marker_begin, lx, pstate,
                           particle_box, origin_a, origin_b, origin_ok, fol_a, fol_b,
                           (unsigned long long*)owner);
        if ((rc = check_launch("batch_spread_claim"))) return rc;
      }
      // per-marker claims only where some particle has no shared-memory box
      if (!particle_box || (lattice_mask & 16))
        hipLaunchKernelGGL(k_batch_spread_prep, dim3(grid_markers), dim3(kIbmBlock), 0, s, cd, hdr,
                           dyn, marker_case, case_marker_begin, lx, origin_a, origin_b, origin_ok,
                           fol_a, fol_b, (unsigned long long*)owner, marker_particle, particle_box);
      if ((rc = check_launch("batch_spread_prep"))) return rc;
      hipLaunchKernelGGL(k_batch_spread_write, dim3(grid_markers), dim3(kIbmBlock), 0, s, cd, hdr,
                         dyn, marker_case, case_marker_begin, lx, lF, lm, origin_a, origin_b,
                         origin_ok, fol_a, fol_b, (const unsigned long long*)owner, fol_tiles);
      if ((rc = check_launch("batch_spread_write"))) return rc;
    }
    auto fused = [&](auto l) {
      using L = decltype(l);
      constexpr int li = L::Q == 13 ? 0 : L::Q == 15 ? 1 : L::Q == 19 ? 2 : 3;
      const int grid = lattice_grid ? lattice_grid[li] : grid_tiles;
      hipLaunchKernelGGL((k_batch_fused<S, L>), dim3(grid), dim3(kBlock), 0, s, cd, tile_case,
                         tile_local, hdr, (const int*)case_left, (const int*)case_par,
                         (const int*)case_done, (E*)pool_a, (E*)pool_b, (const float*)fol_a,
                         (const float*)fol_b, (const uint32_t*)fol_tiles);
      return check_launch("batch_fused");
    };
    // only the lattices that have cases (an empty launch still sweeps its grid)
    if (((lattice_mask & 1) && (rc = fused(D3Q13{}))) || ((lattice_mask & 2) && (rc = fused(D3Q15{}))) ||
        ((lattice_mask & 4) && (rc = fused(D3Q19{}))) || ((lattice_mask & 8) && (rc = fused(D3Q27{}))))
      return rc;
    if (with_ibm) {
      hipLaunchKernelGGL(k_batch_pose, dim3(grid_particles), dim3(kIbmBlock), 0, s, hdr, dyn,
                         particle_case, particle_markers, lag, lx, pstate, traj, traj_cap);
      if ((rc = check_launch("batch_pose"))) return rc;
    }
    hipLaunchKernelGGL(k_batch_advance, dim3(4), dim3(256), 0, s, h, dyn);
    return check_launch("batch_advance");
  });
}

}  // extern "C"
