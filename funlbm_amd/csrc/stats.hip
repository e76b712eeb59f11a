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

// partial layout per block: min f, max f, sum f, min u, max u, sum u,
// min rho, max rho, sum rho (double)
constexpr int kStats = 9;

template <class L, class S, bool FORCE>
__global__ __launch_bounds__(kBlock) void k_stats(
    KParams p, const typename S::elem* __restrict__ src, const float* __restrict__ fol_old,
    const float* __restrict__ fol_new, const uint32_t* __restrict__ fol_tiles,
    const int* __restrict__ ctl, const float* __restrict__ rho_prev, double* __restrict__ part) {
  using T = typename S::T;
  double v[kStats] = {INFINITY, -INFINITY, 0.0, INFINITY, -INFINITY, 0.0, INFINITY, -INFINITY, 0.0};
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
#pragma unroll
    for (int k = 0; k < L::Q; ++k) {
      const double x = (double)(float)f[k];
      v[0] = fmin(v[0], x);
      v[1] = fmax(v[1], x);
      v[2] += x;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double x = (double)(float)u[d];
      v[3] = fmin(v[3], x);
      v[4] = fmax(v[4], x);
      v[5] += x;
    }
    const double r = (double)(float)rho;
    v[6] = fmin(v[6], r);
    v[7] = fmax(v[7], r);
    v[8] += r;
  }
  __shared__ double red[kStats][kBlock];
#pragma unroll
  for (int i = 0; i < kStats; ++i) red[i][threadIdx.x] = v[i];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int i = 0; i < kStats; i += 3) {
        red[i][threadIdx.x] = fmin(red[i][threadIdx.x], red[i][threadIdx.x + s]);
        red[i + 1][threadIdx.x] = fmax(red[i + 1][threadIdx.x], red[i + 1][threadIdx.x + s]);
        red[i + 2][threadIdx.x] += red[i + 2][threadIdx.x + s];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < kStats) part[(int64_t)blockIdx.x * kStats + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_stats_final(const double* __restrict__ part, int nblocks, double count_f,
                              double count_u, double count_r, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double v[kStats] = {INFINITY, -INFINITY, 0.0, INFINITY, -INFINITY, 0.0, INFINITY, -INFINITY, 0.0};
  for (int b = 0; b < nblocks; ++b) {
    for (int i = 0; i < kStats; i += 3) {
      v[i] = fmin(v[i], part[b * kStats + i]);
      v[i + 1] = fmax(v[i + 1], part[b * kStats + i + 1]);
      v[i + 2] += part[b * kStats + i + 2];
    }
  }
  // out: f min, mean, max, u min, mean, max, rho min, mean, max
  out[0] = v[0];
  out[1] = v[2] / count_f;
  out[2] = v[1];
  out[3] = v[3];
  out[4] = v[5] / count_u;
  out[5] = v[4];
  out[6] = v[6];
  out[7] = v[8] / count_r;
  out[8] = v[7];
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
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, s, partials, blocks, (double)p.n * lat->q,
                     (double)p.n * 3, (double)p.n, out);
  return check_launch("flow_stats");
}

}  // extern "C"
