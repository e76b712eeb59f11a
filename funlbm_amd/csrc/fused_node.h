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

// The fused step of a 16-bit state in deviation arithmetic (step_node_dev,
// collide_dev): the populations stay scaled deviations from the load to the
// store, which rounds to nearest or, with p.sr, stochastically with the
// dither of this step's seed mix (mix32 of the step's seed).
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
