// Per-node body of the fused timestep, shared by the single-lattice kernel
// (fluid.hip) and the multi-lattice kernel (batch.hip).
//
// For node `node` of lattice p: pull-stream the post-collision state with the
// face overrides; where FOL_new != 0 add the forcing term with u' from
// FOL_old; moments with FOL_new; zero u on the masked faces; then either
// collide and store (MAT = false) or write the reference's f, u, rho, p
// (MAT = true).  Follows src/funlbm/lbm/base.py:159-175 and the cul_equ that
// opens the next step (:161).
#pragma once

#include "lbm_common.h"

namespace funlbm {

constexpr int kBlock = 256;  // nodes per block; also the tile of the FOL tile table

template <class S>
struct FusedIO {
  const typename S::elem* __restrict__ src;
  typename S::elem* __restrict__ dst;
  const float* __restrict__ fol_old;
  const float* __restrict__ fol_new;
  const float* __restrict__ rho_prev;
  float* __restrict__ rho_next;
  float* __restrict__ out_f;
  float* __restrict__ out_u;
  float* __restrict__ out_rho;
  float* __restrict__ out_p;
};

// Forcing (where FOL_new != 0, with u' from FOL_old) and the moments of a
// node whose pulled populations are in f.
template <class L, class S, bool FORCE>
__device__ __forceinline__ void after_pull(const KParams& p, int64_t node, const Node& c,
                                           const FusedIO<S>& io, bool new_here, bool old_here,
                                           typename S::T* f, typename S::T& rho, typename S::T* u) {
  using T = typename S::T;
  T Fn[3] = {(T)0, (T)0, (T)0};
  if (FORCE) {
#pragma unroll
    for (int d = 0; d < 3; ++d)
      Fn[d] = (T)(new_here ? io.fol_new[(int64_t)d * p.n + node] + p.g[d] : p.g[d]);
    if (Fn[0] != (T)0 || Fn[1] != (T)0 || Fn[2] != (T)0) {
      // u' of the first moment update (old FOL), then cul_equ2 with the new FOL.
      T Fo[3];
#pragma unroll
      for (int d = 0; d < 3; ++d)
        Fo[d] = (T)(old_here ? io.fol_old[(int64_t)d * p.n + node] + p.g[d] : p.g[d]);
      T r1, u1[3];
      moments<L, T, S::fast>(p, c, f, Fo, r1, u1);
      add_forcing<L, T, S::fast>(f, u1, Fn, relax_of<T>(p));
    }
  }
  moments<L, T, S::fast>(p, c, f, Fn, rho, u);
}

// The post-step f, u, rho of a node (the reference's tensors after the step),
// without the collision.  Shared by fused_node and the observers below.
template <class L, class S, bool FORCE>
__device__ __forceinline__ void step_node_dev(const KParams& p, int64_t node, const Node& c, const FusedIO<S>& io,
                                              bool new_here, bool old_here, float* g, float& rho, float* u,
                                              float* drho = nullptr);

// A 16-bit state's nodes take the deviation arithmetic of its fused step, so
// that u and rho come out of sums of the small deviations, not of the
// populations w_k + g/scale (whose float32 rounding would cost ~1e-8 of u).
template <class L, class S, bool FORCE>
__device__ __forceinline__ void step_node(const KParams& p, int64_t node, const Node& c,
                                          const FusedIO<S>& io, bool new_here, bool old_here,
                                          typename S::T* f, typename S::T& rho,
                                          typename S::T* u) {
  if constexpr (S::fast) {
    step_node_dev<L, S, FORCE>(p, node, c, io, new_here, old_here, f, rho, u);
    static_for<0, L::Q>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      f[K] = __builtin_fmaf(f[K], p.inv_scale, wt<float>(L::W[K]));
    });
  } else {
    pull_node<L, S>(p, io.src, c, node, io.rho_prev, f);
    after_pull<L, S, FORCE>(p, node, c, io, new_here, old_here, f, rho, u);
  }
}

// Store of a full wave's 16-bit populations through LDS: each lane writes
// its Q values into the wave's staging area (direction-major, 64 nodes per
// direction), then the wave stores 16-byte chunks (8 consecutive nodes of one
// direction) with one global_store_dwordx4 per lane: ceil(8 Q / 64) store
// instructions for the wave instead of Q two-byte ones (float32 states gain
// nothing from it on the MI355X).  The wave's 64 nodes start at a multiple
// of 64 in one state tile, so every chunk is 16-byte aligned in the
// destination and in the staging area.
template <class L, class S>
__device__ __forceinline__ void store_wave_staged(const KParams& p, const FusedIO<S>& io, int tn, uint32_t l0,
                                                  const typename S::raw* vals) {
  using R = typename S::raw;
  constexpr int Q = L::Q;
  __shared__ __align__(16) R stage[kBlock / 64][Q * 64];
  R* st = stage[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  static_for<0, Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    st[K * 64 + lane] = vals[K];
  });
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int ts = p.ts;
  char* tile = (char*)io.dst + ((int64_t)Q + (int64_t)tn * Q) * ((int64_t)sizeof(R) << ts);
  constexpr int per = 16 / (int)sizeof(R);  // nodes per chunk
  constexpr int chunks = 64 / per;            // chunks per direction
#pragma unroll
  for (int it = 0; it < (chunks * Q + 63) / 64; ++it) {
    const int c = it * 64 + lane;
    if (c < chunks * Q) {
      const int K = c / chunks, part = c % chunks;
      typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
      const u32x4 v = *(const u32x4*)(st + K * 64 + part * per);
      u32x4* d = (u32x4*)(tile + ((int64_t)K << ts) * (int64_t)sizeof(R) + (int64_t)(l0 + part * per) * sizeof(R));
      __builtin_nontemporal_store(v, d);
    }
  }
}


// Collision of a stepped node and the store of its post-collision
// populations into the destination state (and, next to moving walls, its
// density for the next step's wall term).
template <class L, class S>
__device__ __forceinline__ void collide_store(const KParams& p, int64_t node, const Node& c,
                                              const FusedIO<S>& io, typename S::T* f, typename S::T rho,
                                              const typename S::T* u) {
  using T = typename S::T;
  collide<L, T, S::fast>(f, u, rho, relax_of<T>(p), (T*)nullptr);
  if (p.has_moving && io.rho_next) {
    bool mf = false;
#pragma unroll
    for (int face = 0; face < 6; ++face)
      mf |= (p.bc[face] == FUNLBM_BC_WALL_WITH_SPEED) && on_face(p, c, face);
    if (mf) io.rho_next[node] = (float)rho;
  }
  // The state is written once and read back only by the next step, after
  // every other plane has streamed through the caches: non-temporal stores.
  // The wave's 64 nodes lie in one state tile (every caller runs waves of 64
  // consecutive nodes from a multiple of 64): uniform tile base + in-tile
  // byte offset, as in pull_node.
  using R = typename S::raw;
  constexpr int Q = L::Q;
  const int ts = p.ts;
  const uint32_t tmask = (1u << ts) - 1u;
  const int tn = (__builtin_amdgcn_readfirstlane((int)node) & ~63) >> ts;
  const uint32_t lb = ((uint32_t)node & tmask) * (uint32_t)sizeof(R);
  static_for<0, Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    char* base = (char*)io.dst + ((int64_t)Q + (int64_t)tn * Q + K) * ((int64_t)1 << ts) * (int64_t)sizeof(R);
    __builtin_nontemporal_store(S::store_raw(f[K], wt<float>(L::W[K]), p.inv_scale, p.scale), (R*)(base + lb));
  });
}

template <class L, class S, bool FORCE, bool MAT>
__device__ __forceinline__ void fused_node(const KParams& p, int64_t node, const FusedIO<S>& io,
                                           bool new_here, bool old_here) {
  using T = typename S::T;
  const Node c = node_coords(p, node);
  T f[L::Q];
  T rho, u[3];
  step_node<L, S, FORCE>(p, node, c, io, new_here, old_here, f, rho, u);

  if constexpr (MAT) {
    if (io.out_f) {
#pragma unroll
      for (int k = 0; k < L::Q; ++k) io.out_f[(int64_t)k * p.n + node] = (float)f[k];
    }
    if (io.out_u) {
#pragma unroll
      for (int d = 0; d < 3; ++d) io.out_u[(int64_t)d * p.n + node] = (float)u[d];
    }
    if (io.out_rho) io.out_rho[node] = (float)rho;
    if (io.out_p) io.out_p[node] = (float)(rho / (T)3);
  } else {
    collide_store<L, S>(p, node, c, io, f, rho, u);
  }
}

// Pull, forcing and moments of a node of a 16-bit state in deviation
// arithmetic (moments_dev, add_forcing_dev): g holds the scaled deviations
// (f - w_k) scale of the pulled populations after the forcing.
template <class L, class S, bool FORCE>
__device__ __forceinline__ void step_node_dev(const KParams& p, int64_t node, const Node& c, const FusedIO<S>& io,
                                              bool new_here, bool old_here, float* g, float& rho, float* u,
                                              float* drho) {
  using D = DevOf<S>;
  pull_node<L, D>(p, io.src, c, node, io.rho_prev, g);
  float Fn[3] = {0.0f, 0.0f, 0.0f};
  if (FORCE) {
#pragma unroll
    for (int d = 0; d < 3; ++d) Fn[d] = new_here ? io.fol_new[(int64_t)d * p.n + node] + p.g[d] : p.g[d];
    if (Fn[0] != 0.0f || Fn[1] != 0.0f || Fn[2] != 0.0f) {
      float Fo[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) Fo[d] = old_here ? io.fol_old[(int64_t)d * p.n + node] + p.g[d] : p.g[d];
      float r1, u1[3];
      moments_dev<L>(p, c, g, Fo, r1, u1);
      add_forcing_dev<L>(p, g, u1, Fn);
    }
  }
  moments_dev<L>(p, c, g, Fn, rho, u, drho);
}

// Collision of a 16-bit node in deviation arithmetic (collide_dev) and the
// store of its populations, which rounds to nearest or, with p.sr,
// stochastically with the dither of this step's seed mix (mix32 of the step's
// seed).
template <class L, class S>
__device__ __forceinline__ void collide_store_dev(const KParams& p, int64_t node, const Node& c,
                                                  const FusedIO<S>& io, float* g, float rho, float drho,
                                                  const float* u, uint32_t seed_mix) {
  collide_dev<L>(p, g, u, rho, drho);
  if (p.has_moving && io.rho_next) {
    bool mf = false;
#pragma unroll
    for (int face = 0; face < 6; ++face)
      mf |= (p.bc[face] == FUNLBM_BC_WALL_WITH_SPEED) && on_face(p, c, face);
    if (mf) io.rho_next[node] = rho;
  }
  using R = typename S::raw;
  constexpr int Q = L::Q;
  const int ts = p.ts;
  const uint32_t tmask = (1u << ts) - 1u;
  const int tn = (__builtin_amdgcn_readfirstlane((int)node) & ~63) >> ts;
  const uint32_t lb = ((uint32_t)node & tmask) * (uint32_t)sizeof(R);
  // one straight store sequence per rounding mode (a per-direction select
  // would split the stores into blocks and cost an address add each)
  if (p.sr) {
    const uint32_t h = mix32((uint32_t)node ^ seed_mix);
    // two conversions per register (low and high half), no register to clear
    uint32_t pair = 0;
    static_for<0, Q>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      char* base = (char*)io.dst + ((int64_t)Q + (int64_t)tn * Q + K) * ((int64_t)1 << ts) * (int64_t)sizeof(R);
      pair = S::template cvt_sr_half<(K & 1) != 0>(pair, g[K], __umul24(h, dither_mult(K)));
      __builtin_nontemporal_store((R)(K & 1 ? pair >> 16 : pair), (R*)(base + lb));
    });
  } else {
    static_for<0, Q>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      char* base = (char*)io.dst + ((int64_t)Q + (int64_t)tn * Q + K) * ((int64_t)1 << ts) * (int64_t)sizeof(R);
      __builtin_nontemporal_store(S::from_float(g[K]), (R*)(base + lb));
    });
  }
}

// The fused step of a 16-bit state in deviation arithmetic: the populations
// stay scaled deviations from the load to the store.  The collision and store
// are written out here as in collide_store_dev: with the call, the compiler
// merged the multi-lattice kernel's pull paths differently and that kernel
// ran 40 % slower on the MI355X.
template <class L, class S, bool FORCE>
__device__ __forceinline__ void fused_node_dev(const KParams& p, int64_t node, const FusedIO<S>& io,
                                               bool new_here, bool old_here, uint32_t seed_mix) {
  const Node c = node_coords(p, node);
  float g[L::Q];
  float rho, u[3], drho;
  step_node_dev<L, S, FORCE>(p, node, c, io, new_here, old_here, g, rho, u, &drho);
  collide_dev<L>(p, g, u, rho, drho);
  if (p.has_moving && io.rho_next) {
    bool mf = false;
#pragma unroll
    for (int face = 0; face < 6; ++face)
      mf |= (p.bc[face] == FUNLBM_BC_WALL_WITH_SPEED) && on_face(p, c, face);
    if (mf) io.rho_next[node] = rho;
  }
  using R = typename S::raw;
  constexpr int Q = L::Q;
  const int ts = p.ts;
  const uint32_t tmask = (1u << ts) - 1u;
  const int nw = __builtin_amdgcn_readfirstlane((int)node) & ~63;  // the wave's first node
  const int tn = nw >> ts;
  const uint32_t lb = ((uint32_t)node & tmask) * (uint32_t)sizeof(R);
  R vals[Q];
  if (p.sr) {
    const uint32_t h = mix32((uint32_t)node ^ seed_mix);
    uint32_t pair = 0;
    static_for<0, Q>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      pair = S::template cvt_sr_half<(K & 1) != 0>(pair, g[K], __umul24(h, dither_mult(K)));
      vals[K] = (R)(K & 1 ? pair >> 16 : pair);
    });
  } else {
    static_for<0, Q>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      vals[K] = S::from_float(g[K]);
    });
  }
  if ((int64_t)nw + 64 <= p.n) {
    store_wave_staged<L, S>(p, io, tn, (uint32_t)nw & tmask, vals);
  } else {
    // a partial last wave: per-lane stores of the nodes it holds
    static_for<0, Q>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      char* base = (char*)io.dst + ((int64_t)Q + (int64_t)tn * Q + K) * ((int64_t)1 << ts) * (int64_t)sizeof(R);
      __builtin_nontemporal_store(vals[K], (R*)(base + lb));
    });
  }
}

// Statistics of the reference's f, u and rho of one node (the values as
// float32, as the reference's tensors hold them), accumulated per thread in
// float for the extrema and in double for the sums.  The sum of a node's
// populations is its density (the moments' sum), so the mean of f is taken
// as the mean of rho over Q.
constexpr int kStats = 9;  // partial: min f, max f, sum f, min u, max u, sum u, min rho, max rho, sum rho

struct NodeStats {
  float fmin, fmax, umin, umax, rmin, rmax;
  double usum, rsum;
};

__device__ inline void stats_identity(NodeStats& a) {
  a.fmin = a.umin = a.rmin = INFINITY;
  a.fmax = a.umax = a.rmax = -INFINITY;
  a.usum = a.rsum = 0.0;
}

template <int Q>
__device__ inline void stats_add(NodeStats& a, const float* f, const float* u, float rho) {
  float lo = f[0], hi = f[0];
#pragma unroll
  for (int k = 1; k < Q; ++k) {
    lo = fminf(lo, f[k]);
    hi = fmaxf(hi, f[k]);
  }
  a.fmin = fminf(a.fmin, lo);
  a.fmax = fmaxf(a.fmax, hi);
  a.umin = fminf(a.umin, fminf(fminf(u[0], u[1]), u[2]));
  a.umax = fmaxf(a.umax, fmaxf(fmaxf(u[0], u[1]), u[2]));
  a.usum += ((double)u[0] + (double)u[1]) + (double)u[2];
  a.rmin = fminf(a.rmin, rho);
  a.rmax = fmaxf(a.rmax, rho);
  a.rsum += (double)rho;
}

// Wave-wide reduction, result in lane 63: inclusive scans within the rows of
// 16 lanes (row_shr 1, 2, 4, 8), then the row totals across rows
// (row_bcast 15 into rows 1 and 3, row_bcast 31 into rows 2 and 3); lanes
// without a source take the identity.
template <class Op>
__device__ inline float wave_reduce_f(float v, float ident, Op op) {
  const int id = __float_as_int(ident);
  v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x111, 0xf, 0xf, false)));
  v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x112, 0xf, 0xf, false)));
  v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x114, 0xf, 0xf, false)));
  v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x118, 0xf, 0xf, false)));
  v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x142, 0xa, 0xf, false)));
  v = op(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x143, 0xc, 0xf, false)));
  return v;
}

template <int CTRL, int ROW_MASK>
__device__ inline double dpp_shift_f64(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum over the wave (identity 0.0, whose bits are all zero), result in lane 63.
__device__ inline double wave_sum_d(double v) {
  v += dpp_shift_f64<0x111, 0xf>(v);
  v += dpp_shift_f64<0x112, 0xf>(v);
  v += dpp_shift_f64<0x114, 0xf>(v);
  v += dpp_shift_f64<0x118, 0xf>(v);
  v += dpp_shift_f64<0x142, 0xa>(v);
  v += dpp_shift_f64<0x143, 0xc>(v);
  return v;
}

// Block reduction of the per-thread statistics into part[kStats] (call with
// the whole block): DPP within each wave, then the four waves' totals in
// fixed order.
__device__ inline void stats_block_reduce(NodeStats a, double* __restrict__ part) {
  auto mn = [](float x, float y) { return fminf(x, y); };
  auto mx = [](float x, float y) { return fmaxf(x, y); };
  a.fmin = wave_reduce_f(a.fmin, INFINITY, mn);
  a.fmax = wave_reduce_f(a.fmax, -INFINITY, mx);
  a.umin = wave_reduce_f(a.umin, INFINITY, mn);
  a.umax = wave_reduce_f(a.umax, -INFINITY, mx);
  a.rmin = wave_reduce_f(a.rmin, INFINITY, mn);
  a.rmax = wave_reduce_f(a.rmax, -INFINITY, mx);
  a.usum = wave_sum_d(a.usum);
  a.rsum = wave_sum_d(a.rsum);
  constexpr int W = kBlock / 64;
  __shared__ NodeStats wv[W];
  if ((threadIdx.x & 63) == 63) wv[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    NodeStats t = wv[0];
#pragma unroll
    for (int w = 1; w < W; ++w) {
      t.fmin = fminf(t.fmin, wv[w].fmin);
      t.fmax = fmaxf(t.fmax, wv[w].fmax);
      t.umin = fminf(t.umin, wv[w].umin);
      t.umax = fmaxf(t.umax, wv[w].umax);
      t.rmin = fminf(t.rmin, wv[w].rmin);
      t.rmax = fmaxf(t.rmax, wv[w].rmax);
      t.usum += wv[w].usum;
      t.rsum += wv[w].rsum;
    }
    part[0] = t.fmin;
    part[1] = t.fmax;
    part[2] = t.rsum;
    part[3] = t.umin;
    part[4] = t.umax;
    part[5] = t.usum;
    part[6] = t.rmin;
    part[7] = t.rmax;
    part[8] = t.rsum;
  }
}

// Fixed-order tree over the block of double statistics (k_stats_final).
__device__ inline void stats_block_reduce_d(double v[kStats], double* __restrict__ part) {
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
  if (threadIdx.x < kStats) part[threadIdx.x] = red[threadIdx.x][0];
}

// First stage of the reduction of many partials: block g takes partials
// [g c, g c + c), c = ceil(nblocks / gridDim.x), thread-strided in fixed order
// then the block tree, and writes partial g of `out`.
static __global__ __launch_bounds__(kBlock) void k_stats_stage(const double* __restrict__ part, int nblocks,
                                                               double* __restrict__ out) {
  const int c = (nblocks + (int)gridDim.x - 1) / (int)gridDim.x;
  const int b0 = (int)blockIdx.x * c;
  const int b1 = min(nblocks, b0 + c);
  double v[kStats];
#pragma unroll
  for (int i = 0; i < kStats; i += 3) {
    v[i] = INFINITY;
    v[i + 1] = -INFINITY;
    v[i + 2] = 0.0;
  }
  for (int b = b0 + (int)threadIdx.x; b < b1; b += kBlock) {
#pragma unroll
    for (int i = 0; i < kStats; i += 3) {
      v[i] = fmin(v[i], part[(int64_t)b * kStats + i]);
      v[i + 1] = fmax(v[i + 1], part[(int64_t)b * kStats + i + 1]);
      v[i + 2] += part[(int64_t)b * kStats + i + 2];
    }
  }
  stats_block_reduce_d(v, out + (int64_t)blockIdx.x * kStats);
}

// Fixed-order reduction of nblocks partials (one block of kBlock threads:
// thread t takes partials t, t + kBlock, ..., then the block tree) into
// out = f min, mean, max, u min, mean, max, rho min, mean, max.
static __global__ __launch_bounds__(kBlock) void k_stats_final(const double* __restrict__ part, int nblocks,
                                                               double count_f, double count_u, double count_r,
                                                               double* __restrict__ out) {
  double v[kStats];
#pragma unroll
  for (int i = 0; i < kStats; i += 3) {
    v[i] = INFINITY;
    v[i + 1] = -INFINITY;
    v[i + 2] = 0.0;
  }
  for (int b = threadIdx.x; b < nblocks; b += kBlock) {
#pragma unroll
    for (int i = 0; i < kStats; i += 3) {
      v[i] = fmin(v[i], part[(int64_t)b * kStats + i]);
      v[i + 1] = fmax(v[i + 1], part[(int64_t)b * kStats + i + 1]);
      v[i + 2] += part[(int64_t)b * kStats + i + 2];
    }
  }
  __shared__ double tot[kStats];
  stats_block_reduce_d(v, tot);
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = tot[0];
    out[1] = tot[2] / count_f;
    out[2] = tot[1];
    out[3] = tot[3];
    out[4] = tot[5] / count_u;
    out[5] = tot[4];
    out[6] = tot[6];
    out[7] = tot[8] / count_r;
    out[8] = tot[7];
  }
}

// The fused step of one node that also adds the reference's f, u, rho after
// this step's stream and moments (the fields the step collides) to v.
template <class L, class S, bool FORCE>
__device__ __forceinline__ void fused_stats_node(const KParams& p, int64_t node, const FusedIO<S>& io,
                                                 bool new_here, bool old_here, uint32_t seed_mix,
                                                 NodeStats& v) {
  using T = typename S::T;
  const Node c = node_coords(p, node);
  if constexpr (S::fast) {
    float g[L::Q], f[L::Q];
    float rho, u[3], drho;
    step_node_dev<L, S, FORCE>(p, node, c, io, new_here, old_here, g, rho, u, &drho);
    static_for<0, L::Q>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      f[K] = __builtin_fmaf(g[K], p.inv_scale, wt<float>(L::W[K]));
    });
    stats_add<L::Q>(v, f, u, rho);
    collide_store_dev<L, S>(p, node, c, io, g, rho, drho, u, seed_mix);
  } else {
    T f[L::Q], rho, u[3];
    step_node<L, S, FORCE>(p, node, c, io, new_here, old_here, f, rho, u);
    float ff[L::Q], uf[3];
#pragma unroll
    for (int k = 0; k < L::Q; ++k) ff[k] = (float)f[k];
#pragma unroll
    for (int d = 0; d < 3; ++d) uf[d] = (float)u[d];
    stats_add<L::Q>(v, ff, uf, (float)rho);
    collide_store<L, S>(p, node, c, io, f, rho, u);
  }
}

// The fused step of one node: deviation arithmetic for the 16-bit states,
// the reference's operations otherwise.
template <class L, class S, bool FORCE>
__device__ __forceinline__ void fused_step_node(const KParams& p, int64_t node, const FusedIO<S>& io,
                                                bool new_here, bool old_here, uint32_t seed_mix) {
  if constexpr (S::fast)
    fused_node_dev<L, S, FORCE>(p, node, io, new_here, old_here, seed_mix);
  else
    fused_node<L, S, FORCE, false>(p, node, io, new_here, old_here);
}

}  // namespace funlbm
