// Fluid kernels of the D3Qn BGK timestep and their C-ABI entry points.
//
// Layout: populations f[k][x][y][z] (z fastest), one thread per node over a
// flattened 1-D grid of 256-thread blocks, so rows of any length keep every
// lane busy and every access to a direction plane is a coalesced row.
#include <mutex>

#include "capi_util.h"
#include "fused_node.h"
#include "lbm_common.h"

namespace funlbm {


// ---------------------------------------------------------------------------
// Explicit (reference-structured) kernels, float, one pass each.
// ---------------------------------------------------------------------------
template <class L>
__global__ __launch_bounds__(kBlock) void k_collide(KParams p, float* __restrict__ f,
                                                    const float* __restrict__ u,
                                                    const float* __restrict__ rho,
                                                    float* __restrict__ feq) {
  const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (node >= p.n) return;
  float fl[L::Q], e[L::Q];
#pragma unroll
  for (int k = 0; k < L::Q; ++k) fl[k] = f[(int64_t)k * p.n + node];
  const float uu[3] = {u[node], u[p.n + node], u[2 * p.n + node]};
  collide<L, float>(fl, uu, rho[node], relax_of<float>(p), e);
#pragma unroll
  for (int k = 0; k < L::Q; ++k) {
    f[(int64_t)k * p.n + node] = fl[k];
    if (feq) feq[(int64_t)k * p.n + node] = e[k];
  }
}

template <class L, class S>
__global__ __launch_bounds__(kBlock) void k_stream(KParams p, const typename S::elem* __restrict__ src,
                                                   typename S::elem* __restrict__ dst,
                                                   const float* __restrict__ rho,
                                                   const float* __restrict__ feq) {
  using T = typename S::T;
  const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (node >= p.n) return;
  const Node c = node_coords(p, node);
  static_for<0, L::Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    const Src s = source_of<L, K>(p, c);
    if (s.neq != 0) {
      // f = feq + (f_streamed - feq) of the next plane in (float32 state):
      // the streamed value there is the plain roll of x_in - e_k (periodic)
      constexpr int ex = L::E[K][0], ey = L::E[K][1], ez = L::E[K][2];
      const int xin = s.neq > 0 ? 1 : p.nx - 2;
      const int sx = xin - ex;  // in [0, nx) for nx >= 3
      const int sy = (c.y - ey + p.ny) % p.ny;
      const int sz = (c.z - ez + p.nz) % p.nz;
      const int64_t kn = (int64_t)K * p.n;
      const float fin = (float)src[kn + lin(p, sx, sy, sz)];
      const float v = feq[kn + node] + (fin - feq[kn + lin(p, xin, c.y, c.z)]);
      dst[kn + node] = (typename S::elem)v;
      return;
    }
    typename S::elem v = src[(int64_t)s.k * p.n + lin(p, s.x, s.y, s.z)];
    if (s.moving_face >= 0) {
      const float wr = wt<float>(L::W[K]);
      T val = S::load(v, wr, p.inv_scale, p.scale);
      val = val - moving_wall_term<L, K, T>(p, s.moving_face, (T)rho[node]);
      v = S::store(val, wr, p.inv_scale, p.scale);
    }
    dst[(int64_t)K * p.n + node] = v;
  });
}

// NON_EQUILIBRIUM moments: the max of rho (positive floats order as their
// bit patterns), then the face overrides of update_u_rou_boundary.
__global__ __launch_bounds__(kBlock) void k_fill_zero(float* __restrict__ x) { x[0] = 0.0f; }

__global__ __launch_bounds__(kBlock) void k_rho_max(int64_t n, const float* __restrict__ rho,
                                                    unsigned int* __restrict__ out) {
  float m = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    m = fmaxf(m, rho[i]);
  __shared__ float red[kBlock];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, __float_as_uint(red[0]));
}

__global__ __launch_bounds__(kBlock) void k_neq_apply(KParams p, float* __restrict__ u, float* __restrict__ rho,
                                                      const double* __restrict__ profile, float remu,
                                                      const float* __restrict__ rho_max) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;  // (y, z) of a face plane
  if (i >= (int64_t)p.sxy) return;
  const int64_t in0 = i, in1 = (int64_t)p.sxy + i;                                  // x = 0, 1
  const int64_t out0 = (int64_t)(p.nx - 1) * p.sxy + i, out1 = (int64_t)(p.nx - 2) * p.sxy + i;
  if (p.bc[0] == FUNLBM_BC_NON_EQUILIBRIUM) {
    const float uw = (remu / rho_max[0]) / (float)min(p.ny, p.nz);
    u[in0] = (float)((double)uw * profile[i]);
    rho[in0] = rho[in1];
  }
  if (p.bc[1] == FUNLBM_BC_NON_EQUILIBRIUM) {
    rho[out0] = rho[out1];
#pragma unroll
    for (int d = 0; d < 3; ++d) u[(int64_t)d * p.n + out0] = u[(int64_t)d * p.n + out1];
  }
}

template <class L>
__global__ __launch_bounds__(kBlock) void k_moments(KParams p, const float* __restrict__ f,
                                                    const float* __restrict__ fol,
                                                    float* __restrict__ u, float* __restrict__ rho,
                                                    float* __restrict__ pr) {
  const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (node >= p.n) return;
  const Node c = node_coords(p, node);
  float fl[L::Q];
#pragma unroll
  for (int k = 0; k < L::Q; ++k) fl[k] = f[(int64_t)k * p.n + node];
  float F[3] = {p.g[0], p.g[1], p.g[2]};
  if (fol) {
#pragma unroll
    for (int d = 0; d < 3; ++d) F[d] = fol[(int64_t)d * p.n + node] + p.g[d];
  }
  float r, uu[3];
  moments<L, float>(p, c, fl, F, r, uu);
  rho[node] = r;
  if (pr) pr[node] = r / 3.0f;
#pragma unroll
  for (int d = 0; d < 3; ++d) u[(int64_t)d * p.n + node] = uu[d];
}

template <class L>
__global__ __launch_bounds__(kBlock) void k_forcing(KParams p, float* __restrict__ f,
                                                    const float* __restrict__ u,
                                                    const float* __restrict__ fol) {
  const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (node >= p.n) return;
  float F[3] = {p.g[0], p.g[1], p.g[2]};
  if (fol) {
#pragma unroll
    for (int d = 0; d < 3; ++d) F[d] = fol[(int64_t)d * p.n + node] + p.g[d];
  }
  float fl[L::Q];
#pragma unroll
  for (int k = 0; k < L::Q; ++k) fl[k] = f[(int64_t)k * p.n + node];
  const float uu[3] = {u[node], u[p.n + node], u[2 * p.n + node]};
  add_forcing<L, float>(fl, uu, F, relax_of<float>(p));
#pragma unroll
  for (int k = 0; k < L::Q; ++k) f[(int64_t)k * p.n + node] = fl[k];
}

template <class L, class S>
__global__ __launch_bounds__(kBlock) void k_pack(KParams p, const float* __restrict__ f,
                                                 typename S::elem* __restrict__ state) {
  const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (node >= p.n) return;
  static_for<0, L::Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    const float wr = wt<float>(L::W[K]);
    state[state_index<L::Q>(p, K, node)] =
        S::store((typename S::T)f[(int64_t)K * p.n + node], wr, p.inv_scale, p.scale);
  });
}

// Equilibrium populations of (rho, u) computed in the storage's arithmetic
// type (NULL rho = 1, NULL u = 0), written [q][n] (unpadded planes).
template <class L, class S>
__global__ __launch_bounds__(kBlock) void k_equilibrium(KParams p, const float* __restrict__ rho,
                                                        const float* __restrict__ u,
                                                        typename S::elem* __restrict__ state) {
  using T = typename S::T;
  const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (node >= p.n) return;
  const T r = rho ? (T)rho[node] : (T)1;
  T uu[3] = {(T)0, (T)0, (T)0};
  if (u) {
#pragma unroll
    for (int d = 0; d < 3; ++d) uu[d] = (T)u[(int64_t)d * p.n + node];
  }
  T f[L::Q], e[L::Q];
#pragma unroll
  for (int k = 0; k < L::Q; ++k) f[k] = (T)0;
  collide<L, T>(f, uu, r, Relax<T>{(T)1, (T)1, (T)0}, e);
  static_for<0, L::Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    const float wr = wt<float>(L::W[K]);
    state[(int64_t)K * p.n + node] = S::store(e[K], wr, p.inv_scale, p.scale);
  });
}

// The reference's initial state as the fused path enters it: f = rho w with
// rho = 1, the moments of that f (FOL = the body force), then the extra
// collision of LBMD3._init and the first step's cul_equ, both with those same
// moments (src/funlbm/flow/d3/base.py:99-101, src/funlbm/lbm/lbm3d.py:29,
// src/funlbm/lbm/base.py:161).  Computed in the storage's arithmetic type.
template <class L, class S>
__global__ __launch_bounds__(kBlock) void k_init_state(KParams p, typename S::elem* __restrict__ state) {
  using T = typename S::T;
  const int64_t node = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (node >= p.n) return;
  const Node c = node_coords(p, node);
  T f[L::Q];
  static_for<0, L::Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    f[K] = (((T)1 + (T)0) + (T)0 - (T)0) * ((T)1 * wt<T>(L::W[K]));
  });
  const T F[3] = {(T)p.g[0], (T)p.g[1], (T)p.g[2]};
  T rho, u[3];
  moments<L, T>(p, c, f, F, rho, u);
  collide<L, T>(f, u, rho, relax_of<T>(p), (T*)nullptr);
  collide<L, T>(f, u, rho, relax_of<T>(p), (T*)nullptr);
  static_for<0, L::Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    const float wr = wt<float>(L::W[K]);
    state[state_index<L::Q>(p, K, node)] = S::store(f[K], wr, p.inv_scale, p.scale);
  });
}

// ---------------------------------------------------------------------------
// Fused pull step, one node per thread (body in fused_node.h).
// ---------------------------------------------------------------------------
//
// The parameter block is read through the kernel-argument segment pointer (it
// is the first argument, at offset 0): a runtime index into it (the moving
// wall's face) is then a load from that segment, where with the by-value
// argument the compiler may copy the whole block to private memory.
template <class L, class S, bool FORCE, bool MAT>
__device__ __forceinline__ void fused_kernel_body(
    const typename S::elem* __restrict__ src, typename S::elem* __restrict__ dst,
    const float* __restrict__ fol_old, const float* __restrict__ fol_new,
    const uint32_t* __restrict__ fol_tiles, const int* __restrict__ ctl,
    const float* __restrict__ rho_prev, float* __restrict__ rho_next, float* __restrict__ out_f,
    float* __restrict__ out_u, float* __restrict__ out_rho, float* __restrict__ out_p) {
  const KParams& p = *(const KParams*)__builtin_amdgcn_kernarg_segment_ptr();
  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const int64_t node = (int64_t)b * kBlock + threadIdx.x;
  if (node >= p.n) return;
  bool new_here = fol_new != nullptr, old_here = fol_old != nullptr;
  if (FORCE && fol_tiles) {
    const uint32_t entry = fol_tiles[(uint32_t)node >> 8], stamp = (uint32_t)ctl[0];
    new_here = new_here && entry == stamp;
    old_here = old_here && (entry == stamp || entry == stamp - 1u);
  }
  FusedIO<S> io{src, dst, fol_old, fol_new, rho_prev, rho_next, out_f, out_u, out_rho, out_p};
  if constexpr (MAT)
    fused_node<L, S, FORCE, true>(p, node, io, new_here, old_here);
  else
    fused_step_node<L, S, FORCE>(p, node, io, new_here, old_here, p.seed_mix);
}

#define FUNLBM_FUSED_ARGS                                                                              \
  KParams p_arg, const typename S::elem *__restrict__ src, typename S::elem *__restrict__ dst,        \
      const float *__restrict__ fol_old, const float *__restrict__ fol_new,                            \
      const uint32_t *__restrict__ fol_tiles, const int *__restrict__ ctl,                             \
      const float *__restrict__ rho_prev, float *__restrict__ rho_next, float *__restrict__ out_f,     \
      float *__restrict__ out_u, float *__restrict__ out_rho, float *__restrict__ out_p

template <class L, class S, bool FORCE, bool MAT>
__global__ __launch_bounds__(kBlock) void k_fused(FUNLBM_FUSED_ARGS) {
  (void)p_arg;
  fused_kernel_body<L, S, FORCE, MAT>(src, dst, fol_old, fol_new, fol_tiles, ctl, rho_prev, rho_next, out_f,
                                      out_u, out_rho, out_p);
}

// The time step proper of the float32 and 16-bit states: at least seven
// waves per SIMD (at most 72 VGPRs).  Left alone the compiler gives the
// D3Q27 16-bit step 74 (six waves); with the bound it fits 63.
template <class L, class S, bool FORCE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_fused_step(
    FUNLBM_FUSED_ARGS) {
  (void)p_arg;
  fused_kernel_body<L, S, FORCE, false>(src, dst, fol_old, fol_new, fol_tiles, ctl, rho_prev, rho_next, out_f,
                                        out_u, out_rho, out_p);
}

constexpr int kStatsStage = 256;  // first-stage blocks of the statistics' final reduction

// The fused step that also reduces the reference's f, u, rho after this
// step's stream and moments (what funlbm_flow_stats reads from the step's
// source state) into one partial per block: the statistics of a logged step
// without a second pass over the state.  Block b (after the XCD remap) takes
// nodes [256 b, 256 b + 256) and writes partial b.
template <class L, class S, bool FORCE>
__global__ __launch_bounds__(kBlock) void k_fused_stats(FUNLBM_FUSED_ARGS, double* __restrict__ part) {
  (void)p_arg;
  const KParams& p = *(const KParams*)__builtin_amdgcn_kernarg_segment_ptr();
  const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
  const int64_t node = (int64_t)b * kBlock + threadIdx.x;
  NodeStats v;
  stats_identity(v);
  if (node < p.n) {
    bool new_here = fol_new != nullptr, old_here = fol_old != nullptr;
    if (FORCE && fol_tiles) {
      const uint32_t entry = fol_tiles[(uint32_t)node >> 8], stamp = (uint32_t)ctl[0];
      new_here = new_here && entry == stamp;
      old_here = old_here && (entry == stamp || entry == stamp - 1u);
    }
    FusedIO<S> io{src, dst, fol_old, fol_new, rho_prev, rho_next, nullptr, nullptr, nullptr, nullptr};
    fused_stats_node<L, S, FORCE>(p, node, io, new_here, old_here, p.seed_mix, v);
  }
  stats_block_reduce(v, part + (int64_t)b * kStats);
}

// Plain copy, one 16-byte element per thread (the fastest plain form measured
// on the MI355X: about 6.1 TB/s read + write on 2 GiB).
__global__ __launch_bounds__(kBlock) void k_copy(const uint4* __restrict__ src,
                                                 uint4* __restrict__ dst, int64_t n16) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n16) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
static int grid_for(int64_t n) { return (int)((n + kBlock - 1) / kBlock); }

}  // namespace funlbm

using namespace funlbm;

extern "C" {

int funlbm_abi_version(void) { return FUNLBM_ABI_VERSION; }

const char* funlbm_last_error(void) { return last_error_message(); }

int funlbm_state_bytes(const funlbm_lattice* lat, int64_t* bytes) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (!bytes) return set_error(FUNLBM_EINVAL, "bytes is NULL");
  *bytes = state_elems(lat->q, p.n, p.ts) * (int64_t)storage_elem_bytes(lat->storage);
  return 0;
}

int funlbm_collide(const funlbm_lattice* lat, float* f, const float* u, const float* rho,
                   float* feq, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (!f || !u || !rho) return set_error(FUNLBM_EINVAL, "collide: NULL array");
  hipStream_t s = (hipStream_t)stream;
  return dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    hipLaunchKernelGGL(k_collide<L>, dim3(grid_for(p.n)), dim3(kBlock), 0, s, p, f, u, rho, feq);
    return check_launch("collide");
  });
}

int funlbm_stream(const funlbm_lattice* lat, const void* f_src, void* f_dst, const float* rho,
                  const float* feq, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (!f_src || !f_dst) return set_error(FUNLBM_EINVAL, "stream: NULL array");
  if (f_src == f_dst) return set_error(FUNLBM_EINVAL, "stream: source and destination alias");
  if (p.has_moving && !rho) return set_error(FUNLBM_EINVAL, "stream: moving wall needs rho");
  if (p.has_neq && (!feq || lat->storage != FUNLBM_STORE_F32))
    return set_error(FUNLBM_EINVAL, "stream: NON_EQUILIBRIUM faces need feq and float32 storage");
  hipStream_t s = (hipStream_t)stream;
  return dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    return dispatch_storage(lat->storage, [&](auto st) {
      using S = decltype(st);
      using E = typename S::elem;
      hipLaunchKernelGGL((k_stream<L, S>), dim3(grid_for(p.n)), dim3(kBlock), 0, s, p,
                         (const E*)f_src, (E*)f_dst, rho, feq);
      return check_launch("stream");
    });
  });
}

int funlbm_moments(const funlbm_lattice* lat, const float* f, const float* fol, float* u,
                   float* rho, float* pr, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (!f || !u || !rho) return set_error(FUNLBM_EINVAL, "moments: NULL array");
  hipStream_t s = (hipStream_t)stream;
  return dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    hipLaunchKernelGGL(k_moments<L>, dim3(grid_for(p.n)), dim3(kBlock), 0, s, p, f, fol, u, rho, pr);
    return check_launch("moments");
  });
}

int funlbm_neq_moments(const funlbm_lattice* lat, float* u, float* rho, const double* profile, float remu,
                       float* work, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (!p.has_neq) return 0;
  if (!u || !rho || !work) return set_error(FUNLBM_EINVAL, "neq_moments: NULL array");
  const bool inlet = p.bc[0] == FUNLBM_BC_NON_EQUILIBRIUM;
  if (inlet && !profile) return set_error(FUNLBM_EINVAL, "neq_moments: an input NEQ face needs the profile");
  hipStream_t s = (hipStream_t)stream;
  if (inlet) {
    hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(1), 0, s, work);
    if ((rc = check_launch("neq_moments"))) return rc;
    const int blocks = (int)std::min<int64_t>((p.n + kBlock - 1) / kBlock, 2048);
    hipLaunchKernelGGL(k_rho_max, dim3(blocks), dim3(kBlock), 0, s, p.n, (const float*)rho, (unsigned int*)work);
    if ((rc = check_launch("neq_moments"))) return rc;
  }
  hipLaunchKernelGGL(k_neq_apply, dim3(grid_for(p.sxy)), dim3(kBlock), 0, s, p, u, rho, profile, remu,
                     (const float*)work);
  return check_launch("neq_moments");
}

int funlbm_forcing(const funlbm_lattice* lat, float* f, const float* u, const float* fol,
                   void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (!f || !u) return set_error(FUNLBM_EINVAL, "forcing: NULL array");
  hipStream_t s = (hipStream_t)stream;
  return dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    hipLaunchKernelGGL(k_forcing<L>, dim3(grid_for(p.n)), dim3(kBlock), 0, s, p, f, u, fol);
    return check_launch("forcing");
  });
}

int funlbm_pack(const funlbm_lattice* lat, const float* f, void* state, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (!f || !state) return set_error(FUNLBM_EINVAL, "pack: NULL array");
  hipStream_t s = (hipStream_t)stream;
  return dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    return dispatch_storage(lat->storage, [&](auto st) {
      using S = decltype(st);
      hipLaunchKernelGGL((k_pack<L, S>), dim3(grid_for(p.n)), dim3(kBlock), 0, s, p, f,
                         (typename S::elem*)state);
      return check_launch("pack");
    });
  });
}

int funlbm_equilibrium_state(const funlbm_lattice* lat, const float* rho, const float* u,
                             void* state, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (!state) return set_error(FUNLBM_EINVAL, "equilibrium: NULL state");
  hipStream_t s = (hipStream_t)stream;
  return dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    return dispatch_storage(lat->storage, [&](auto st) {
      using S = decltype(st);
      hipLaunchKernelGGL((k_equilibrium<L, S>), dim3(grid_for(p.n)), dim3(kBlock), 0, s, p, rho, u,
                         (typename S::elem*)state);
      return check_launch("equilibrium");
    });
  });
}

int funlbm_init_state(const funlbm_lattice* lat, void* state, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if ((rc = reject_neq(p, "init_state"))) return rc;
  if (!state) return set_error(FUNLBM_EINVAL, "init_state: NULL state");
  hipStream_t s = (hipStream_t)stream;
  return dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    return dispatch_storage(lat->storage, [&](auto st) {
      using S = decltype(st);
      hipLaunchKernelGGL((k_init_state<L, S>), dim3(grid_for(p.n)), dim3(kBlock), 0, s, p,
                         (typename S::elem*)state);
      return check_launch("init_state");
    });
  });
}

static int fused_launch(const funlbm_lattice* lat, const void* src, void* dst, const float* fol_old,
                        const float* fol_new, const uint32_t* fol_tiles, const int32_t* ctl,
                        const float* rho_prev, float* rho_next, float* out_f, float* out_u,
                        float* out_rho, float* out_p, bool mat, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if ((rc = reject_neq(p, mat ? "materialize" : "fused_step"))) return rc;
  if (!src || (!mat && !dst)) return set_error(FUNLBM_EINVAL, "fused: NULL state");
  if (!mat && src == dst) return set_error(FUNLBM_EINVAL, "fused: source and destination alias");
  if (p.has_moving && !rho_prev)
    return set_error(FUNLBM_EINVAL, "fused: moving wall needs the previous face density");
  if (fol_tiles && !ctl) return set_error(FUNLBM_EINVAL, "fused: a FOL tile table needs ctl");
  const bool force = fol_old || fol_new || p.has_body;
  hipStream_t s = (hipStream_t)stream;
  return dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    return dispatch_storage(lat->storage, [&](auto st) {
      using S = decltype(st);
      using E = typename S::elem;
      auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid_for(p.n)), dim3(kBlock), 0, s, p, (const E*)src, (E*)dst,
                           fol_old, fol_new, fol_tiles, ctl, rho_prev, rho_next, out_f, out_u, out_rho,
                           out_p);
        return check_launch(mat ? "materialize" : "fused_step");
      };
      if (mat) return force ? go(k_fused<L, S, true, true>) : go(k_fused<L, S, false, true>);
      if constexpr (S::code == FUNLBM_STORE_F64)
        return force ? go(k_fused<L, S, true, false>) : go(k_fused<L, S, false, false>);
      else
        return force ? go(k_fused_step<L, S, true>) : go(k_fused_step<L, S, false>);
    });
  });
}

int funlbm_fused_step(const funlbm_lattice* lat, const void* state_src, void* state_dst,
                      const float* fol_old, const float* fol_new, const uint32_t* fol_tiles,
                      const int32_t* ctl, const float* rho_prev, float* rho_next, void* stream) {
  return fused_launch(lat, state_src, state_dst, fol_old, fol_new, fol_tiles, ctl, rho_prev, rho_next,
                      nullptr, nullptr, nullptr, nullptr, false, stream);
}

// The logged fused step with its block partials in `partials` (n_partials
// entries of kStats doubles); `who` names the entry point in errors.
static int fused_stats_launch(const funlbm_lattice* lat, const void* state_src, void* state_dst,
                              const float* fol_old, const float* fol_new, const uint32_t* fol_tiles,
                              const int32_t* ctl, const float* rho_prev, float* rho_next, double* partials,
                              int64_t n_partials, double* out, void* stream, const char* who) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if ((rc = reject_neq(p, who))) return rc;
  if (!state_src || !state_dst) return set_error(FUNLBM_EINVAL, "%s: NULL state", who);
  if (state_src == state_dst) return set_error(FUNLBM_EINVAL, "%s: source and destination alias", who);
  if (!partials || !out) return set_error(FUNLBM_EINVAL, "%s: NULL statistics array", who);
  if (p.has_moving && !rho_prev)
    return set_error(FUNLBM_EINVAL, "%s: moving wall needs the previous face density", who);
  if (fol_tiles && !ctl) return set_error(FUNLBM_EINVAL, "%s: a FOL tile table needs ctl", who);
  const int blocks = grid_for(p.n);
  const int stage = std::min(blocks, kStatsStage);
  if (n_partials < (int64_t)blocks + stage)
    return set_error(FUNLBM_EINVAL, "%s: %lld partials given, %d needed (one per 256 nodes, and %d more)", who,
                     (long long)n_partials, blocks + stage, stage);
  const bool force = fol_old || fol_new || p.has_body;
  hipStream_t s = (hipStream_t)stream;
  rc = dispatch_lattice(lat->q, [&](auto l) {
    using L = decltype(l);
    return dispatch_storage(lat->storage, [&](auto st) {
      using S = decltype(st);
      using E = typename S::elem;
      auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), 0, s, p, (const E*)state_src, (E*)state_dst, fol_old,
                           fol_new, fol_tiles, ctl, rho_prev, rho_next, nullptr, nullptr, nullptr, nullptr,
                           partials);
        return check_launch(who);
      };
      return force ? go(k_fused_stats<L, S, true>) : go(k_fused_stats<L, S, false>);
    });
  });
  if (rc) return rc;
  // many partials: a first stage over `stage` blocks, then one block
  double* second = partials + (int64_t)blocks * kStats;
  hipLaunchKernelGGL(k_stats_stage, dim3(stage), dim3(kBlock), 0, s, partials, blocks, second);
  if ((rc = check_launch(who))) return rc;
  hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(kBlock), 0, s, second, stage, (double)p.n * lat->q,
                     (double)p.n * 3, (double)p.n, out);
  return check_launch(who);
}

// ---------------------------------------------------------------------------
// Scratch of funlbm_fused_step_stats_noalloc: the block partials of a logged
// step of up to kNoAllocNodes nodes (one partial per 256 nodes and
// kStatsStage more, kStats doubles each: 4.5 MiB), in a file-scope device
// array that each GPU holds for the life of the process.  Host code never
// passes the array's name to a kernel (in host code it is a stand-in, not a
// device address): the address comes from hipGetSymbolAddress, once per
// device with that device current, and is kept per device.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocNodes = (int64_t)1 << 24;  // 256^3
constexpr int64_t kNoAllocPartials = kNoAllocNodes / kBlock + kStatsStage;
static __device__ __align__(256) double scratch[kNoAllocPartials * kStats];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static double* g_scratch[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on its first use.
static int scratch_address(double** out, const char* who) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error((int)e, "%s: hipGetDevice: %s", who, hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return set_error(FUNLBM_EINVAL, "%s: device %d beyond %d", who, dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return set_error((int)e, "%s: hipGetSymbolAddress of the scratch array: %s", who, hipGetErrorString(e));
    if (!ptr) return set_error(FUNLBM_EFAILED, "%s: the scratch array has no device address", who);
    g_scratch[dev] = (double*)ptr;
  }
  *out = g_scratch[dev];
  return 0;
}

int funlbm_fused_step_stats(const funlbm_lattice* lat, const void* state_src, void* state_dst,
                            const float* fol_old, const float* fol_new, const uint32_t* fol_tiles,
                            const int32_t* ctl, const float* rho_prev, float* rho_next, double* partials,
                            int32_t n_partials, double* out, void* stream) {
  return fused_stats_launch(lat, state_src, state_dst, fol_old, fol_new, fol_tiles, ctl, rho_prev, rho_next,
                            partials, n_partials, out, stream, "fused_step_stats");
}

int funlbm_fused_step_stats_noalloc(const funlbm_lattice* lat, const void* state_src, void* state_dst,
                                    const float* fol_old, const float* fol_new, const uint32_t* fol_tiles,
                                    const int32_t* ctl, const float* rho_prev, float* rho_next, double* out,
                                    void* stream) {
  const char* who = "fused_step_stats_noalloc";
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (p.n > kNoAllocNodes)
    return set_error(FUNLBM_ETOOBIG, "%s: %lld nodes exceed the %lld-node maximum of its scratch array", who,
                     (long long)p.n, (long long)kNoAllocNodes);
  double* partials = nullptr;
  if ((rc = scratch_address(&partials, who))) return rc;
  return fused_stats_launch(lat, state_src, state_dst, fol_old, fol_new, fol_tiles, ctl, rho_prev, rho_next,
                            partials, kNoAllocPartials, out, stream, who);
}

int funlbm_materialize(const funlbm_lattice* lat, const void* state_src, const float* fol_old,
                       const float* fol_new, const uint32_t* fol_tiles, const int32_t* ctl,
                       const float* rho_prev, float* f, float* u, float* rho, float* pr,
                       void* stream) {
  return fused_launch(lat, state_src, nullptr, fol_old, fol_new, fol_tiles, ctl, rho_prev, nullptr, f,
                      u, rho, pr, true, stream);
}

int funlbm_copy_calibrate(const void* src, void* dst, int64_t n_bytes, void* stream) {
  if (!src || !dst || n_bytes <= 0 || (n_bytes % 16) != 0)
    return set_error(FUNLBM_EINVAL, "copy: bad arguments");
  const int64_t n16 = n_bytes / 16;
  if ((n16 + kBlock - 1) / kBlock > 0x7fffffffLL) return set_error(FUNLBM_ETOOBIG, "copy: too large");
  const int blocks = (int)((n16 + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(kBlock), 0, (hipStream_t)stream,
                     (const uint4*)src, (uint4*)dst, n16);
  return check_launch("copy");
}

}  // extern "C"
