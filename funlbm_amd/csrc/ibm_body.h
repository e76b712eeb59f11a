// Immersed-boundary bodies shared by the single-lattice kernels (ibm.hip)
// and the multi-lattice kernels (batch.hip).
//
// Per marker (16 lanes, four markers per wavefront; lane (j, k) of a marker
// holds the stencil nodes rl + (i, j, k), i = 0..3, so that four consecutive
// lanes touch four consecutive nodes):
//   interp_marker       interpolation + direct forcing
//   spread_prep_marker  clear the old stencil, claim the owner of each node
//   spread_write_marker the owner writes w lF lm
// Per particle (one 256-thread block):
//   wall_particle       wall spring
//   forces_particle     force/torque reduction and rigid-body update (double)
//   pose_particle       rotation and marker positions
#pragma once

#include "lbm_common.h"

namespace funlbm {

constexpr int kWave = 64;
constexpr int kIbmBlock = 256;
constexpr int kMarkerLanes = 16;                        // lanes of one marker
constexpr int kMarkersPerWave = kWave / kMarkerLanes;   // 4
constexpr int kMarkersPerBlock = kIbmBlock / kMarkerLanes;  // 16
constexpr float kPiF = 3.14159265358979323846f;

// Stencil of a marker (src/funlbm/lbm/lbm3d.py:32-44): rl = max(floor(x) - 1, 0),
// rr = min(rl + 4, size) with size = the config size (node count - 1).
// rl is also capped at size before the integer conversion: a marker past the
// upper face has an empty stencil either way (rr = size <= rl), and the cap
// keeps lo + 3 far from integer overflow for any input, NaN and inf included
// (fmaxf(NaN, 0) = 0).
struct Stencil {
  int lo[3], hi[3];
};

__device__ inline Stencil stencil_of(const KParams& p, const float x[3]) {
  Stencil s;
  const int size[3] = {p.nx - 1, p.ny - 1, p.nz - 1};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float rl = floorf(x[d]) - 1.0f;
    rl = fmaxf(rl, 0.0f);
    rl = fminf(rl, (float)size[d]);
    float rr = fminf(rl + 4.0f, (float)size[d]);
    s.lo[d] = (int)rl;
    s.hi[d] = (int)rr;
  }
  return s;
}

// Sum over the 16 lanes of one marker (xor butterfly 8, 4, 2, 1 on lane
// 4j + k: j's bits first, then k's).
template <class T>
__device__ inline T group_sum(T v) {
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, kMarkerLanes);
  return v;
}

// Sum over a marker's 64 stencil nodes, v[i] being node (i, j, k) of the
// lane.  The tree is the xor butterfly 32, 16, 8, 4, 2, 1 over node index
// 16i + 4j + k: its first two levels give (v0 + v2) + (v1 + v3) in each lane,
// the other four are group_sum.
template <class T>
__device__ inline T stencil_sum(const T v[4]) {
  return group_sum((v[0] + v[2]) + (v[1] + v[3]));
}

// Lane geometry of a marker: node (x0 + i, y, z) of each layer i, which of
// them lie in the stencil, and their normalised weights.  The cosine kernel
// (src/funlbm/lbm/lbm3d.py:46-59): phi(r) = (1 + cos(|r pi/2|))/4 per axis
// (or the Peskin kernel, delta_1d), the node's weight (t_x t_y) t_z where all
// |r_d| <= 2, else 0, r = node - x; the weights of the stencil are
// normalised to sum 1.
// Per-axis kernel value at offset r (|r| <= 2 is tested by the caller):
// the reference's cosine kernel, or Peskin's 4-point kernel
// (C. S. Peskin, Acta Numerica 11 (2002) 479-517, eq. 6.27), whose values at
// the four nodes around a point sum to 1 and whose first moment vanishes.
__device__ inline float delta_1d(const KParams& p, float r) {
  if (p.delta == FUNLBM_DELTA_PESKIN4) {
    const float a = fabsf(r);
    if (a <= 1.0f) return (3.0f - 2.0f * a + sqrtf(1.0f + 4.0f * a - 4.0f * a * a)) / 8.0f;
    return (5.0f - 2.0f * a - sqrtf(fmaxf(-7.0f + 12.0f * a - 4.0f * a * a, 0.0f))) / 8.0f;
  }
  return (1.0f + cosf(fabsf(r * kPiF / 2.0f))) / 4.0f;
}

struct LaneW {
  int x0, y, z;
  unsigned valid;  // bit i: node i inside the stencil
  float w[4];      // normalised weights (0 outside)
  float wsum;      // sum of the normalised weights of the stencil
};

// Call with the marker's 16 lanes together (it shuffles and sums within them).
__device__ inline LaneW lane_weight(const KParams& p, const float x[3], int sub) {
  LaneW lw;
  const Stencil s = stencil_of(p, x);
  const int j = sub >> 2, k = sub & 3;
  lw.x0 = s.lo[0];
  lw.y = s.lo[1] + j;
  lw.z = s.lo[2] + k;
  const bool vyz = lw.y < s.hi[1] && lw.z < s.hi[2];
  // unnormalised weight per node: (t_x t_y) t_z, t_y and t_z shared by the layers
  const float ry = (float)lw.y - x[1], rz = (float)lw.z - x[2];
  const bool inyz = fabsf(ry) <= 2.0f && fabsf(rz) <= 2.0f;
  // The marker needs 12 kernel values: t_x of its four layers, t_y of its
  // four y offsets and t_z of its four z offsets.  Lane 4a + o evaluates
  // axis a's value at offset o (lanes 12-15 a spare z one), from the same
  // operands as the lane that uses it, and each lane reads its six values
  // from those lanes: one kernel evaluation per lane instead of six.
  const int ax = sub >> 2, o = sub & 3;
  const float xa = ax == 0 ? x[0] : ax == 1 ? x[1] : x[2];
  const int la = ax == 0 ? s.lo[0] : ax == 1 ? s.lo[1] : s.lo[2];
  const float ta = delta_1d(p, (float)(la + o) - xa);
  const float ty = __shfl(ta, 4 + j, kMarkerLanes);
  const float tz = __shfl(ta, 8 + k, kMarkerLanes);
  lw.valid = 0;
  float w[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const bool v = vyz && lw.x0 + i < s.hi[0];
    lw.valid |= (unsigned)v << i;
    const float rx = (float)(lw.x0 + i) - x[0];
    const float tx = __shfl(ta, i, kMarkerLanes);
    w[i] = (v && inyz && fabsf(rx) <= 2.0f) ? (tx * ty) * tz : 0.0f;
  }
  const float ws = stencil_sum(w);
  if (ws > 0.0f) {
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = w[i] / ws;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) lw.w[i] = w[i];
  lw.wsum = stencil_sum(w);
  return lw;
}

// Particle state layout (double[P][FUNLBM_PSTRIDE]).
__device__ inline double* pst(double* ps, int i) { return ps + (int64_t)i * FUNLBM_PSTRIDE; }
__device__ inline const double* pst(const double* ps, int i) {
  return ps + (int64_t)i * FUNLBM_PSTRIDE;
}

// ---------------------------------------------------------------------------
// Moments of node (x, y, z) as the interpolation sees them: the post-stream
// populations pulled out of the fused state, with FOL_old as the forcing
// (a 16-bit state in its deviation arithmetic, as its fused step).
// ---------------------------------------------------------------------------
template <class L, class S>
__device__ __forceinline__ void state_node_moments(const KParams& p,
                                                   const typename S::elem* __restrict__ state,
                                                   const float* __restrict__ fol_old,
                                                   const float* __restrict__ rho_face, int x, int y,
                                                   int z, float& rr, float uu[3]) {
  using T = typename S::T;
  const int64_t node = lin(p, x, y, z);
  const Node c{x, y, z};
  if constexpr (S::fast) {
    float g[L::Q];
    pull_node<L, DevOf<S>, false>(p, state, c, node, rho_face, g);
    float F[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) F[d] = fol_old ? fol_old[(int64_t)d * p.n + node] + p.g[d] : p.g[d];
    moments_dev<L>(p, c, g, F, rr, uu);
  } else {
    T f[L::Q];
    pull_node<L, S, false>(p, state, c, node, rho_face, f);
    T F[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
      F[d] = (T)(fol_old ? fol_old[(int64_t)d * p.n + node] + p.g[d] : p.g[d]);
    T r, u[3];
    moments<L, T, S::fast>(p, c, f, F, r, u);
    rr = (float)r;
#pragma unroll
    for (int d = 0; d < 3; ++d) uu[d] = (float)u[d];
  }
}

// ---------------------------------------------------------------------------
// Moment cache of a particle's box.  The markers of a particle lie within R
// of its centre cx (R: the ceiling of its largest body-frame marker distance),
// so a stencil node n of a marker x, with
// floor(x) - 1 <= n <= floor(x) + 2 and floor(cx) - R <= floor(x) <=
// floor(cx) + R, lies in the cube [o, o + S)^3 with o = floor(cx) - (R + 1)
// and S = 2R + 4.  One pass (box_node) stores (rho, ux, uy, uz) of every cube
// node inside the lattice, computed as state_node_moments computes it; the
// interpolation then reads each node's moments once per marker instead of
// pulling Q populations per covering marker (about 20 markers per node on a
// marker spacing of one).  A marker with a node outside the cube (a stencil
// clamped at a face, a centre that is not finite, markers not yet placed) is
// left to the per-node interpolation, so the result is the same bit for bit
// either way.
// Box row of a particle, int32[3]: {cache offset in nodes, R (-1: no box),
// first 256-node chunk of the box in the chunk list}; a box that does not fit
// the cache's `cap` nodes is not used.
// ---------------------------------------------------------------------------
constexpr int kBoxChunk = 256;

struct Box {
  int o[3];
  int S;
  int off;
  bool ok;
};

__device__ inline Box box_of(const double* __restrict__ P, const int* __restrict__ row, int cap) {
  Box b;
  const int R = row[1];
  b.off = row[0];
  b.S = 2 * R + 4;
  // R <= 512 keeps S^3 in int; the box must lie inside the cache
  b.ok = R >= 0 && R <= 512 && b.off >= 0 &&
         (int64_t)b.off + (int64_t)b.S * b.S * b.S <= (int64_t)cap;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double c = floor(P[FUNLBM_P_CX + d]);
    const bool fin = c > -1048576.0 && c < 1048576.0;  // false for NaN
    b.ok = b.ok && fin;
    b.o[d] = fin ? (int)c - (R + 1) : 0;
  }
  return b;
}

// Cube node `local` of the box of chunk c (one thread per node): moments into
// cache[off + local].  The chunk list names each chunk's particle.
template <class L, class S>
__device__ inline void box_node(const KParams& p, const typename S::elem* __restrict__ state,
                                const float* __restrict__ fol_old, const float* __restrict__ rho_face,
                                const double* __restrict__ P, const int* __restrict__ row, int chunk,
                                float4* __restrict__ cache, int cap) {
  const Box b = box_of(P, row, cap);
  if (!b.ok) return;
  const int local = (chunk - row[2]) * kBoxChunk + (int)threadIdx.x;
  if (local >= b.S * b.S * b.S) return;
  const int k = local % b.S, t = local / b.S;
  const int j = t % b.S, i = t / b.S;
  const int x = b.o[0] + i, y = b.o[1] + j, z = b.o[2] + k;
  if (x < 0 || x >= p.nx || y < 0 || y >= p.ny || z < 0 || z >= p.nz) return;
  float rr, uu[3];
  state_node_moments<L, S>(p, state, fol_old, rho_face, x, y, z, rr, uu);
  cache[b.off + local] = make_float4(rr, uu[0], uu[1], uu[2]);
}

// ---------------------------------------------------------------------------
// Interpolation + direct forcing (src/funlbm/lbm/lbm3d.py:66-100):
//   s_u = sum_nodes sum_d u_d w (the three components summed into one scalar),
//   s_rho = sum w rho, lu = cu + cw x (lx - cx), lF = 2 s_rho (lu - s_u)/dt.
// interp_marker: u', rho' from the explicit fields u / rho, or (FROM_STATE)
// pulled out of the fused state per node (state_node_moments).
// interp_marker_boxed: from the particle's box cache; a marker with a stencil
// node outside its box is appended to miss_list ([0] = count) instead, for
// interp_marker to finish.  Both give the same bits.  Call with the 16 lanes
// of the marker (sub = lane within them).
// ---------------------------------------------------------------------------
__device__ inline void interp_finish(int m, int sub, const float x[3], const double* __restrict__ P,
                                     float su, float sr, float wsum, float dt, bool check_sum,
                                     float* __restrict__ lu, float* __restrict__ lF,
                                     int* __restrict__ fail_flag) {
  if (sub != 0) return;
  if (check_sum && fabsf(wsum - 1.0f) > 1e-3f) atomicOr(fail_flag, 1);
  const double d[3] = {(double)x[0] - P[FUNLBM_P_CX], (double)x[1] - P[FUNLBM_P_CX + 1],
                       (double)x[2] - P[FUNLBM_P_CX + 2]};
  const double* w = P + FUNLBM_P_CW;
  const double cr[3] = {w[1] * d[2] - w[2] * d[1], w[2] * d[0] - w[0] * d[2],
                        w[0] * d[1] - w[1] * d[0]};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float l = (float)(P[FUNLBM_P_CU + a] + cr[a]);
    lu[3 * m + a] = l;
    lF[3 * m + a] = ((2.0f * sr) * (l - su)) / dt;
  }
}

template <class L, class S, bool FROM_STATE>
__device__ inline void interp_marker(const KParams& p, int m, int sub,
                                     const typename S::elem* __restrict__ state,
                                     const float* __restrict__ fol_old,
                                     const float* __restrict__ rho_face,
                                     const float* __restrict__ ufield,
                                     const float* __restrict__ rhofield,
                                     const float* __restrict__ lx, const double* __restrict__ P,
                                     float dt, bool check_sum, float* __restrict__ lu,
                                     float* __restrict__ lF, int* __restrict__ fail_flag) {
  const float x[3] = {lx[3 * m], lx[3 * m + 1], lx[3 * m + 2]};
  const LaneW lw = lane_weight(p, x, sub);
  float su4[4], sr4[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    su4[k] = 0.0f;
    sr4[k] = 0.0f;
  }
  // one copy of the per-node work, looped over the valid layers
#pragma unroll 1
  for (unsigned left = lw.valid; left; left &= left - 1) {
    const int k = __builtin_ctz(left);
    const int xx = lw.x0 + k;
    float uu[3], rr;
    if constexpr (FROM_STATE) {
      state_node_moments<L, S>(p, state, fol_old, rho_face, xx, lw.y, lw.z, rr, uu);
    } else {
      const int64_t node = lin(p, xx, lw.y, lw.z);
      rr = rhofield[node];
#pragma unroll
      for (int d = 0; d < 3; ++d) uu[d] = ufield[(int64_t)d * p.n + node];
    }
    const float w = k == 0 ? lw.w[0] : k == 1 ? lw.w[1] : k == 2 ? lw.w[2] : lw.w[3];
    const float a = (uu[0] * w + uu[1] * w) + uu[2] * w, r = rr * w;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q == k) {
        su4[q] = a;
        sr4[q] = r;
      }
    }
  }
  interp_finish(m, sub, x, P, stencil_sum(su4), stencil_sum(sr4), lw.wsum, dt, check_sum, lu, lF,
                fail_flag);
}

__device__ inline void interp_marker_boxed(const KParams& p, int m, int sub,
                                           const float* __restrict__ lx,
                                           const double* __restrict__ P, float dt, bool check_sum,
                                           float* __restrict__ lu, float* __restrict__ lF,
                                           int* __restrict__ fail_flag,
                                           const int* __restrict__ box_row,
                                           const float4* __restrict__ cache, int cap,
                                           int* __restrict__ miss_list) {
  const float x[3] = {lx[3 * m], lx[3 * m + 1], lx[3 * m + 2]};
  const LaneW lw = lane_weight(p, x, sub);
  const Box b = box_of(P, box_row, cap);
  const int i0 = lw.x0 - b.o[0], j = lw.y - b.o[1], k = lw.z - b.o[2];
  const bool in_yz = b.ok && (unsigned)j < (unsigned)b.S && (unsigned)k < (unsigned)b.S;
  bool miss = false;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    miss |= ((lw.valid >> i) & 1u) && !(in_yz && (unsigned)(i0 + i) < (unsigned)b.S);
  // the marker's 16 lanes decide together
  const int g = (int)(threadIdx.x & (kWave - 1)) / kMarkerLanes;
  if ((__ballot(miss) >> (kMarkerLanes * g)) & 0xffffull) {
    if (sub == 0) miss_list[1 + atomicAdd(&miss_list[0], 1)] = m;
    return;
  }
  float su4[4], sr4[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    su4[i] = 0.0f;
    sr4[i] = 0.0f;
    if (!((lw.valid >> i) & 1u)) continue;
    const float4 v = cache[b.off + ((i0 + i) * b.S + j) * b.S + k];
    const float w = lw.w[i];
    su4[i] = (v.y * w + v.z * w) + v.w * w;
    sr4[i] = v.x * w;
  }
  interp_finish(m, sub, x, P, stencil_sum(su4), stencil_sum(sr4), lw.wsum, dt, check_sum, lu, lF,
                fail_flag);
}

// ---------------------------------------------------------------------------
// Wall spring (src/funlbm/lbm/lbm3d.py:121-167) for the markers [b, e) of one
// particle.  Extreme markers are the first arg-min / arg-max per axis (torch's
// tie rule); d = max(k0 (1 - |coord - bound|/(2 dx)), 0), bounds
// [0,0,0,size]; the six rows lF[ext_j] + n_j d_j are gathered before any is
// written, then written in order j = 0..5 so a duplicate index keeps the last
// row.  Call with the whole block.
// ---------------------------------------------------------------------------
__device__ inline void wall_particle(const KParams& p, int b, int e, const float* __restrict__ lx,
                                     float* __restrict__ lF, float dx) {
  __shared__ float sv[6][kIbmBlock];
  __shared__ int si[6][kIbmBlock];
  float bv[6];
  int bi[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    bv[j] = (j < 3) ? INFINITY : -INFINITY;
    bi[j] = 0x7fffffff;
  }
  for (int m = b + threadIdx.x; m < e; m += kIbmBlock) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const float v = lx[3 * m + (j % 3)];
      const bool better = (j < 3) ? (v < bv[j] || (v == bv[j] && m < bi[j]))
                                  : (v > bv[j] || (v == bv[j] && m < bi[j]));
      if (better) {
        bv[j] = v;
        bi[j] = m;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    sv[j][threadIdx.x] = bv[j];
    si[j][threadIdx.x] = bi[j];
  }
  __syncthreads();
  for (int s = kIbmBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        const float v = sv[j][threadIdx.x + s];
        const int i = si[j][threadIdx.x + s];
        const float cv = sv[j][threadIdx.x];
        const int ci = si[j][threadIdx.x];
        const bool better =
            (j < 3) ? (v < cv || (v == cv && i < ci)) : (v > cv || (v == cv && i < ci));
        if (better) {
          sv[j][threadIdx.x] = v;
          si[j][threadIdx.x] = i;
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && e > b) {
    const float bounds[6] = {0.0f, 0.0f, 0.0f, (float)(p.nx - 1), (float)(p.ny - 1), (float)(p.nz - 1)};
    const float normals[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
    float dist[6];
    float dmax = 0.0f;
    int ext[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      // every comparison fails on NaN coordinates; stay inside [b, e)
      ext[j] = (si[j][0] >= b && si[j][0] < e) ? si[j][0] : b;
      const float c = lx[3 * ext[j] + (j % 3)];
      float d = 300.0f * (1.0f - fabsf(c - bounds[j]) / (2.0f * dx));
      d = fmaxf(d, 0.0f);
      dist[j] = d;
      dmax = fmaxf(dmax, d);
    }
    if (dmax != 0.0f) {
      float rows[6][3];
#pragma unroll
      for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int a = 0; a < 3; ++a) rows[j][a] = lF[3 * ext[j] + a] + normals[j][a] * dist[j];
#pragma unroll
      for (int j = 0; j < 6; ++j)
#pragma unroll
        for (int a = 0; a < 3; ++a) lF[3 * ext[j] + a] = rows[j][a];
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Rigid-body update (src/funlbm/particle/base/base.py:120-151) in double:
// cF = sum(-lF lm) + (1 - rho_p/rho_f) V g; cu += cF/V dt; cx += cu dt;
// cT = -sum((lx - cx) x lF) lm; cw += 0.1 cT dt / I.  Fixed-order tree
// reductions over the block.  Call with the whole block.
// ---------------------------------------------------------------------------
__device__ inline void forces_particle(int b, int e, const float* __restrict__ lx,
                                       const float* __restrict__ lF, const float* __restrict__ lm,
                                       double* __restrict__ P, double dt, double rho_f) {
  __shared__ double red[6][kIbmBlock];
  __shared__ double cxn[3];
  double fs[3] = {0, 0, 0};
  for (int m = b + threadIdx.x; m < e; m += kIbmBlock)
#pragma unroll
    for (int a = 0; a < 3; ++a) fs[a] += -(double)lF[3 * m + a] * (double)lm[m];
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 3; ++a) red[a][threadIdx.x] = fs[a];
  __syncthreads();
  for (int s = kIbmBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int a = 0; a < 3; ++a) red[a][threadIdx.x] += red[a][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double V = P[FUNLBM_P_MASS];
    const double buoy = (1.0 - P[FUNLBM_P_RHO] / rho_f) * V;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double F = red[a][0] + buoy * P[FUNLBM_P_GRAV + a];
      P[FUNLBM_P_CF + a] = F;
      P[FUNLBM_P_CU + a] = P[FUNLBM_P_CU + a] + F / V * dt;
      P[FUNLBM_P_CX + a] = P[FUNLBM_P_CX + a] + P[FUNLBM_P_CU + a] * dt;
      cxn[a] = P[FUNLBM_P_CX + a];
    }
  }
  __syncthreads();
  double ts[3] = {0, 0, 0};
  for (int m = b + threadIdx.x; m < e; m += kIbmBlock) {
    const double r[3] = {(double)lx[3 * m] - cxn[0], (double)lx[3 * m + 1] - cxn[1],
                         (double)lx[3 * m + 2] - cxn[2]};
    const double f[3] = {lF[3 * m], lF[3 * m + 1], lF[3 * m + 2]};
    const double w = lm[m];
    ts[0] += (r[1] * f[2] - r[2] * f[1]) * w;
    ts[1] += (r[2] * f[0] - r[0] * f[2]) * w;
    ts[2] += (r[0] * f[1] - r[1] * f[0]) * w;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) red[3 + a][threadIdx.x] = ts[a];
  __syncthreads();
  for (int s = kIbmBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int a = 0; a < 3; ++a) red[3 + a][threadIdx.x] += red[3 + a][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double T = -red[3 + a][0];
      P[FUNLBM_P_CT + a] = T;
      P[FUNLBM_P_CW + a] = P[FUNLBM_P_CW + a] + 0.1 * T * dt / P[FUNLBM_P_INERTIA + a];
    }
  }
  __syncthreads();
}

// Rotation matrix of a rotation vector (Rodrigues), double.
__device__ inline void rotvec_matrix(const double v[3], double R[9]) {
  const double th2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const double th = sqrt(th2);
  double s, c1;  // sin(th)/th and (1 - cos(th))/th^2
  if (th < 1e-4) {
    s = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
    c1 = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
  } else {
    s = sin(th) / th;
    c1 = (1.0 - cos(th)) / th2;
  }
  const double x = v[0], y = v[1], z = v[2];
  R[0] = 1.0 - c1 * (y * y + z * z);
  R[1] = -s * z + c1 * x * y;
  R[2] = s * y + c1 * x * z;
  R[3] = s * z + c1 * x * y;
  R[4] = 1.0 - c1 * (x * x + z * z);
  R[5] = -s * x + c1 * y * z;
  R[6] = -s * y + c1 * x * z;
  R[7] = s * x + c1 * y * z;
  R[8] = 1.0 - c1 * (x * x + y * y);
}

// Pose update (src/funlbm/particle/base/base.py:153-156,
// src/funlbm/particle/coord.py:73-108): angle += cw (no dt), R = rotvec(angle),
// lx = float32(R lag) + float32(cx).  Call with the whole block.
__device__ inline void pose_particle(int b, int e, const float* __restrict__ lag,
                                     float* __restrict__ lx, double* __restrict__ P, bool advance) {
  __shared__ double R[9];
  __shared__ float cx[3];
  if (threadIdx.x == 0) {
    double v[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (advance) P[FUNLBM_P_ANGLE + a] = P[FUNLBM_P_ANGLE + a] + P[FUNLBM_P_CW + a];
      v[a] = P[FUNLBM_P_ANGLE + a];
      cx[a] = (float)P[FUNLBM_P_CX + a];
    }
    double r[9];
    rotvec_matrix(v, r);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      R[i] = r[i];
      P[FUNLBM_P_ROT + i] = r[i];
    }
  }
  __syncthreads();
  for (int m = b + threadIdx.x; m < e; m += kIbmBlock) {
    const double q[3] = {lag[3 * m], lag[3 * m + 1], lag[3 * m + 2]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double v = R[3 * a] * q[0] + R[3 * a + 1] * q[1] + R[3 * a + 2] * q[2];
      lx[3 * m + a] = (float)v + cx[a];
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Spreading (src/funlbm/lbm/lbm3d.py:102-119): FOL := 0, then for every marker
// in order FOL[stencil] = w lF lm, overwriting.  A node ends up with the value
// of the highest-numbered marker whose stencil box covers it.
//   prep   zero the stencil this buffer received from the marker last time
//          and raise owner[node] to (stamp, marker) with an integer atomicMax
//   write  the owning marker of each node writes w lF lm and stamps the tile
// The stamp advances by one per spreading, so owner[] never needs clearing.
// A marker whose stencil origin is the one this buffer received it at clears
// nothing: the stencil is the same node set, and every node of it is claimed
// and written again.  The four markers of a wavefront are consecutive and
// their stencils overlap:
// prep leaves a node that a later marker of the wave (of the same lattice)
// also claims, or also clears, to that marker, so the wave issues each claim
// and each clear once.  The atomicMax result is unchanged (the later key is
// the larger one).
// ---------------------------------------------------------------------------
struct PrepPeer {
  int lo[3];  // stencil origin (claims)
  int ol[3];  // last stencil origin of this buffer (clears)
  int flags;  // bit 0: active, bit 1: has a stencil to clear
  int tag;    // lattice of the marker
};

__device__ inline bool in_box4(const int lo[3], int x, int y, int z, int ex, int ey, int ez) {
  // [lo, min(lo + 4, e)) per axis
  return x >= lo[0] && x < min(lo[0] + 4, ex) && y >= lo[1] && y < min(lo[1] + 4, ey) && z >= lo[2] &&
         z < min(lo[2] + 4, ez);
}

// Call with every lane of the wave (inactive markers included: active =
// false), the wave's four markers consecutive.  peers: this wave's four
// entries of block-shared memory.
__device__ inline void spread_prep_marker(const KParams& p, bool active, int tag, int m, int key_id,
                                          int sub, const float* __restrict__ lx,
                                          const int* __restrict__ clear_origin,
                                          float* __restrict__ fol,
                                          unsigned long long* __restrict__ owner, uint32_t stamp,
                                          PrepPeer* peers) {
  const int g = (int)(threadIdx.x & (kWave - 1)) / kMarkerLanes;
  const int j = sub >> 2, k = sub & 3;
  Stencil s{};
  int ol[3] = {0, 0, 0};
  bool clear = active && clear_origin != nullptr;
  if (active) {
    const float x[3] = {lx[3 * m], lx[3 * m + 1], lx[3 * m + 2]};
    s = stencil_of(p, x);
  }
  if (clear) {
#pragma unroll
    for (int d = 0; d < 3; ++d) ol[d] = clear_origin[3 * m + d];
    // an unchanged stencil needs no clearing: each of its nodes is claimed
    // again and written by its owner in this spreading
    clear = ol[0] != s.lo[0] || ol[1] != s.lo[1] || ol[2] != s.lo[2];
  }
  if (sub == 0) {
    PrepPeer& e = peers[g];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      e.lo[d] = s.lo[d];
      e.ol[d] = ol[d];
    }
    e.flags = (active ? 1 : 0) | (clear ? 2 : 0);
    e.tag = tag;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (!active) return;
  const int size[3] = {p.nx - 1, p.ny - 1, p.nz - 1};
  if (clear) {
    const int oy = ol[1] + j, oz = ol[2] + k;
    if (oy < p.ny && oz < p.nz) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int ox = ol[0] + i;
        if (ox >= p.nx) break;
        bool later = false;
        for (int q = g + 1; q < kMarkersPerWave; ++q) {
          const PrepPeer& e = peers[q];
          later |= (e.flags & 2) && e.tag == tag && in_box4(e.ol, ox, oy, oz, p.nx, p.ny, p.nz);
        }
        if (later) continue;
        const int64_t node = lin(p, ox, oy, oz);
#pragma unroll
        for (int d = 0; d < 3; ++d) fol[(int64_t)d * p.n + node] = 0.0f;
      }
    }
  }
  const int ny_ = s.lo[1] + j, nz_ = s.lo[2] + k;
  if (ny_ < s.hi[1] && nz_ < s.hi[2]) {
    const unsigned long long key = ((unsigned long long)stamp << 32) | (unsigned long long)(unsigned)key_id;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int nx_ = s.lo[0] + i;
      if (nx_ >= s.hi[0]) break;
      bool later = false;
      for (int q = g + 1; q < kMarkersPerWave; ++q) {
        const PrepPeer& e = peers[q];
        later |= (e.flags & 1) && e.tag == tag && in_box4(e.lo, nx_, ny_, nz_, size[0], size[1], size[2]);
      }
      if (later) continue;
      atomicMax(&owner[lin(p, nx_, ny_, nz_)], key);
    }
  }
}

// ---------------------------------------------------------------------------
// Claims of one particle's markers through its box, in shared memory.  Each
// marker raises own[node] (the box's nodes, shared memory) to its index in
// the particle with a shared-memory atomicMax: the particle's owner of every
// node it covers, a segmented maximum over its markers.  Then each owned node
// takes one global claim (stamp, key0 + owner); a node two particles cover
// gets the larger key, as with one claim per marker.  A stencil node outside
// the box (the centre moved in this step's rigid-body update, a centre that
// is not finite) is claimed globally at once.  Moved stencils are cleared as
// in spread_prep_marker.  Particles whose box exceeds kClaimNodes nodes, or
// that have none, go through spread_prep_marker.  Call with the whole block.
// ---------------------------------------------------------------------------
constexpr int kClaimNodes = 16384;  // int32 per node: 64 KiB of shared memory
constexpr int kClaimBlock = 1024;   // threads of a claiming block: 64 markers per pass

__device__ inline bool claims_in_box(const int* __restrict__ row) {
  const int R = row[1];
  if (R < 0 || R > 10) return false;
  const int S = 2 * R + 4;
  return S * S * S <= kClaimNodes;
}

__device__ inline void spread_claim_particle(const KParams& p, int b, int e, int key0,
                                             const float* __restrict__ lx,
                                             const double* __restrict__ P,
                                             const int* __restrict__ row,
                                             const int* __restrict__ clear_origin,
                                             float* __restrict__ fol,
                                             unsigned long long* __restrict__ owner,
                                             uint32_t stamp, int* own) {
  const Box box = box_of(P, row, 0x7fffffff);
  const int S = box.S, n3 = S * S * S;
  for (int t = threadIdx.x; t < n3; t += blockDim.x) own[t] = -1;
  __syncthreads();
  const unsigned long long hi = (unsigned long long)stamp << 32;
  const int sub = threadIdx.x & (kMarkerLanes - 1);
  const int j = sub >> 2, k = sub & 3;
  const int per_pass = (int)blockDim.x / kMarkerLanes;
  for (int m0 = b; m0 < e; m0 += per_pass) {
    const int m = m0 + (int)threadIdx.x / kMarkerLanes;
    if (m >= e) continue;
    const float x[3] = {lx[3 * m], lx[3 * m + 1], lx[3 * m + 2]};
    const Stencil s = stencil_of(p, x);
    if (clear_origin) {
      const int ol[3] = {clear_origin[3 * m], clear_origin[3 * m + 1], clear_origin[3 * m + 2]};
      const int oy = ol[1] + j, oz = ol[2] + k;
      if ((ol[0] != s.lo[0] || ol[1] != s.lo[1] || ol[2] != s.lo[2]) && oy < p.ny && oz < p.nz) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (ol[0] + i >= p.nx) break;
          const int64_t node = lin(p, ol[0] + i, oy, oz);
#pragma unroll
          for (int d = 0; d < 3; ++d) fol[(int64_t)d * p.n + node] = 0.0f;
        }
      }
    }
    const int ny_ = s.lo[1] + j, nz_ = s.lo[2] + k;
    if (ny_ >= s.hi[1] || nz_ >= s.hi[2]) continue;
    const int lj = ny_ - box.o[1], lk = nz_ - box.o[2];
    const bool in_yz = box.ok && (unsigned)lj < (unsigned)S && (unsigned)lk < (unsigned)S;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int nx_ = s.lo[0] + i;
      if (nx_ >= s.hi[0]) break;
      const int li = nx_ - box.o[0];
      if (in_yz && (unsigned)li < (unsigned)S)
        atomicMax(&own[(li * S + lj) * S + lk], m - b);
      else
        atomicMax(&owner[lin(p, nx_, ny_, nz_)], hi | (unsigned)(key0 + (m - b)));
    }
  }
  __syncthreads();
  if (!box.ok) return;
  for (int t = threadIdx.x; t < n3; t += blockDim.x) {
    const int w = own[t];
    if (w < 0) continue;
    const int lk = t % S, r = t / S;
    const int lj = r % S, li = r / S;
    atomicMax(&owner[lin(p, box.o[0] + li, box.o[1] + lj, box.o[2] + lk)],
              hi | (unsigned)(key0 + w));
  }
}

__device__ inline void spread_write_marker(const KParams& p, int m, int key_id, int sub,
                                           const float* __restrict__ lx,
                                           const float* __restrict__ lF,
                                           const float* __restrict__ lm, float* __restrict__ fol,
                                           const unsigned long long* __restrict__ owner,
                                           int* __restrict__ origin, uint32_t* __restrict__ tiles,
                                           uint32_t stamp) {
  const float x[3] = {lx[3 * m], lx[3 * m + 1], lx[3 * m + 2]};
  const LaneW lw = lane_weight(p, x, sub);
  if (sub == 0 && origin) {
    const Stencil s = stencil_of(p, x);
    origin[3 * m] = s.lo[0];
    origin[3 * m + 1] = s.lo[1];
    origin[3 * m + 2] = s.lo[2];
  }
  if (!lw.valid) return;
  const unsigned long long key = ((unsigned long long)stamp << 32) | (unsigned long long)(unsigned)key_id;
  const float f[3] = {lF[3 * m], lF[3 * m + 1], lF[3 * m + 2]};
  const float mm = lm[m];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (!((lw.valid >> i) & 1u)) continue;
    const int64_t node = lin(p, lw.x0 + i, lw.y, lw.z);
    if (owner[node] != key) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d) fol[(int64_t)d * p.n + node] = (lw.w[i] * f[d]) * mm;
    if (tiles) tiles[(uint32_t)node >> 8] = stamp;
  }
}

}  // namespace funlbm
