// min / mean / max of the reference's f, u and rho after a fused step, read
// straight from the state (the per-step statistics of
// src/funlbm/lbm/base.py:105-144 and src/funlbm/flow/base.py:183-206) without
// materialising the fields.  One pass: every block reduces its nodes, a
// second single-block kernel reduces the block partials in a fixed order, so
// the result is deterministic.
#include "capi_util.h"
#include "fused_node.h"
#include "lbm_common.h"

namespace funlbm {

// partial layout per block (kStats, fused_node.h): min f, max f, sum f,
// min u, max u, sum u, min rho, max rho, sum rho (double)

template <class L, class S, bool FORCE>
__global__ __launch_bounds__(kBlock) void k_stats(
    KParams p, const typename S::elem* __restrict__ src, const float* __restrict__ fol_old,
    const float* __restrict__ fol_new, const uint32_t* __restrict__ fol_tiles,
    const int* __restrict__ ctl, const float* __restrict__ rho_prev, double* __restrict__ part) {
  using T = typename S::T;
  NodeStats v;
  stats_identity(v);
  for (int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x; node < p.n;
       node += (int64_t)gridDim.x * kBlock) {
    bool new_here = fol_new != nullptr, old_here = fol_old != nullptr;
    if (FORCE && fol_tiles) {
      const uint32_t entry = fol_tiles[(uint32_t)node >> 8], stamp = (uint32_t)ctl[0];
      new_here = new_here && entry == stamp;
      old_here = old_here && (entry == stamp || entry == stamp - 1u);
    }
    FusedIO<S> io{src, nullptr, fol_old, fol_new, rho_prev, nullptr, nullptr, nullptr, nullptr, nullptr};
    const Node c = node_coords(p, node);
    T f[L::Q], rho, u[3];
    step_node<L, S, FORCE>(p, node, c, io, new_here, old_here, f, rho, u);
    // values as float32, as the reference's tensors hold them
    float ff[L::Q], uf[3];
#pragma unroll
    for (int k = 0; k < L::Q; ++k) ff[k] = (float)f[k];
#pragma unroll
    for (int d = 0; d < 3; ++d) uf[d] = (float)u[d];
    stats_add<L::Q>(v, ff, uf, (float)rho);
  }
  stats_block_reduce(v, part + (int64_t)blockIdx.x * kStats);
}

constexpr int kWave64 = 64;

// Global mass and momentum of the reference's rho and u after a fused step,
// in double: sum rho, sum rho u_d over the nodes (rho and u as the
// reference's float32 tensors hold them).  Per block: four double sums
// (DPP within each wave, then the waves in fixed order) into part[4 b].
template <class L, class S, bool FORCE>
__global__ __launch_bounds__(kBlock) void k_conservation(
    KParams p, const typename S::elem* __restrict__ src, const float* __restrict__ fol_old,
    const float* __restrict__ fol_new, const uint32_t* __restrict__ fol_tiles,
    const int* __restrict__ ctl, const float* __restrict__ rho_prev, double* __restrict__ part) {
  using T = typename S::T;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x; node < p.n;
       node += (int64_t)gridDim.x * kBlock) {
    bool new_here = fol_new != nullptr, old_here = fol_old != nullptr;
    if (FORCE && fol_tiles) {
      const uint32_t entry = fol_tiles[(uint32_t)node >> 8], stamp = (uint32_t)ctl[0];
      new_here = new_here && entry == stamp;
      old_here = old_here && (entry == stamp || entry == stamp - 1u);
    }
    FusedIO<S> io{src, nullptr, fol_old, fol_new, rho_prev, nullptr, nullptr, nullptr, nullptr, nullptr};
    const Node c = node_coords(p, node);
    T f[L::Q], rho, u[3];
    step_node<L, S, FORCE>(p, node, c, io, new_here, old_here, f, rho, u);
    const double r = (double)(float)rho;
    acc[0] += r;
#pragma unroll
    for (int d = 0; d < 3; ++d) acc[1 + d] += r * (double)(float)u[d];
  }
  __shared__ double red[kBlock / kWave64][4];
  const int w = threadIdx.x / kWave64, lane = threadIdx.x & (kWave64 - 1);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double v = wave_sum_d(acc[i]);  // the wave's total lands in lane 63
    if (lane == kWave64 - 1) red[w][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int k = 0; k < kBlock / kWave64; ++k) s += red[k][threadIdx.x];
    part[(int64_t)blockIdx.x * 4 + threadIdx.x] = s;
  }
}

// Fixed-order sum of nblocks partials of four doubles (one block).
static __global__ __launch_bounds__(kBlock) void k_conservation_final(const double* __restrict__ part,
                                                                      int nblocks,
                                                                      double* __restrict__ out) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += kBlock) {
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] += part[(int64_t)b * 4 + i];
  }
  __shared__ double red[kBlock / kWave64][4];
  const int w = threadIdx.x / kWave64, lane = threadIdx.x & (kWave64 - 1);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double v = wave_sum_d(acc[i]);
    if (lane == kWave64 - 1) red[w][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int k = 0; k < kBlock / kWave64; ++k) s += red[k][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

}  // namespace funlbm

using namespace funlbm;

extern "C" {

int funlbm_flow_stats(const funlbm_lattice* lat, const void* state_src, const float* fol_old,
                      const float* fol_new, const uint32_t* fol_tiles, const int32_t* ctl,
                      const float* rho_prev, double* partials, int32_t n_partials, double* out,
                      void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if ((rc = reject_neq(p, "flow_stats"))) return rc;
  if (!state_src || !partials || !out) return set_error(FUNLBM_EINVAL, "flow_stats: NULL array");
  if (n_partials < 1) return set_error(FUNLBM_EINVAL, "flow_stats: need at least one partial");
  if (p.has_moving && !rho_prev)
    return set_error(FUNLBM_EINVAL, "flow_stats: moving wall needs the previous face density");
  if (fol_tiles && !ctl) return set_error(FUNLBM_EINVAL, "flow_stats: a FOL tile table needs ctl");
  const int64_t want = (p.n + kBlock - 1) / kBlock;
  const int blocks = (int)std::min<int64_t>(want, n_partials);
  const bool force = fol_old || fol_new || p.has_body;
  hipStream_t s = (hipStream_t)stream;
  rc = dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    return dispatch_storage(lat->storage, [&](auto st) {
      using S = decltype(st);
      using E = typename S::elem;
      if (force)
        hipLaunchKernelGGL((k_stats<L, S, true>), dim3(blocks), dim3(kBlock), 0, s, p, (const E*)state_src,
                           fol_old, fol_new, fol_tiles, ctl, rho_prev, partials);
      else
        hipLaunchKernelGGL((k_stats<L, S, false>), dim3(blocks), dim3(kBlock), 0, s, p, (const E*)state_src,
                           fol_old, fol_new, fol_tiles, ctl, rho_prev, partials);
      return check_launch("flow_stats");
    });
  });
  if (rc) return rc;
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(kBlock), 0, s, partials, blocks, (double)p.n * lat->q,
                     (double)p.n * 3, (double)p.n, out);
  return check_launch("flow_stats");
}

int funlbm_flow_conservation(const funlbm_lattice* lat, const void* state_src, const float* fol_old,
                             const float* fol_new, const uint32_t* fol_tiles, const int32_t* ctl,
                             const float* rho_prev, double* partials, int32_t n_partials, double* out,
                             void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if ((rc = reject_neq(p, "flow_conservation"))) return rc;
  if (!state_src || !partials || !out) return set_error(FUNLBM_EINVAL, "flow_conservation: NULL array");
  if (n_partials < 1) return set_error(FUNLBM_EINVAL, "flow_conservation: need at least one partial");
  if (p.has_moving && !rho_prev)
    return set_error(FUNLBM_EINVAL, "flow_conservation: moving wall needs the previous face density");
  if (fol_tiles && !ctl) return set_error(FUNLBM_EINVAL, "flow_conservation: a FOL tile table needs ctl");
  const int64_t want = (p.n + kBlock - 1) / kBlock;
  const int blocks = (int)std::min<int64_t>(want, n_partials);
  const bool force = fol_old || fol_new || p.has_body;
  hipStream_t s = (hipStream_t)stream;
  rc = dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    return dispatch_storage(lat->storage, [&](auto st) {
      using S = decltype(st);
      using E = typename S::elem;
      if (force)
        hipLaunchKernelGGL((k_conservation<L, S, true>), dim3(blocks), dim3(kBlock), 0, s, p,
                           (const E*)state_src, fol_old, fol_new, fol_tiles, ctl, rho_prev, partials);
      else
        hipLaunchKernelGGL((k_conservation<L, S, false>), dim3(blocks), dim3(kBlock), 0, s, p,
                           (const E*)state_src, fol_old, fol_new, fol_tiles, ctl, rho_prev, partials);
      return check_launch("flow_conservation");
    });
  });
  if (rc) return rc;
  hipLaunchKernelGGL(k_conservation_final, dim3(1), dim3(kBlock), 0, s, partials, blocks, out);
  return check_launch("flow_conservation");
}

}  // extern "C"
