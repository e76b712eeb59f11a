// Lattice tables, storage formats and the per-node arithmetic shared by every
// fluid and IBM kernel of the library.  gfx950 only.
//
// The per-node arithmetic follows the operation order of the reference so that
// float results agree to rounding (src/funlbm/flow/d3/base.py:103-159):
// no fused multiply-add is formed (the library is built with -ffp-contract=off),
// divisions stay divisions and sums run in direction order.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/funlbm.h"

namespace funlbm {

// Compile-time loop: f(std::integral_constant<int, I>) for I in [B, E).
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// ---------------------------------------------------------------------------
// Lattice tables.  Direction order, velocities and weights are the reference's
// (src/funlbm/flow/d3/d3q13.py:7-44, d3q15.py:7-48, d3q19.py:7-60,
// d3q27.py:7-77); REV[k] is the index of -e_k (the reference builds it from a
// string map, d3q19.py:55-60).
// ---------------------------------------------------------------------------
struct D3Q13 {
  static constexpr int Q = 13;
  static constexpr int E[13][3] = {{0, 0, 0},  {1, 1, 0},  {1, -1, 0},  {1, 0, 1},  {1, 0, -1},
                                   {0, 1, 1},  {0, 1, -1}, {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1},
                                   {-1, 0, 1}, {0, -1, -1}, {0, -1, 1}};
  static constexpr double W[13] = {1.0 / 2,  1.0 / 24, 1.0 / 24, 1.0 / 24, 1.0 / 24,
                                   1.0 / 24, 1.0 / 24, 1.0 / 24, 1.0 / 24, 1.0 / 24,
                                   1.0 / 24, 1.0 / 24, 1.0 / 24};
  static constexpr int REV[13] = {0, 7, 8, 9, 10, 11, 12, 1, 2, 3, 4, 5, 6};
};

struct D3Q15 {
  static constexpr int Q = 15;
  static constexpr int E[15][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},  {0, -1, 0},
                                   {0, 0, 1},   {0, 0, -1},  {1, 1, 1},   {-1, 1, 1}, {1, -1, 1},
                                   {-1, -1, 1}, {1, 1, -1},  {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
  static constexpr double W[15] = {2.0 / 9,  1.0 / 9,  1.0 / 9,  1.0 / 9,  1.0 / 9,
                                   1.0 / 9,  1.0 / 9,  1.0 / 72, 1.0 / 72, 1.0 / 72,
                                   1.0 / 72, 1.0 / 72, 1.0 / 72, 1.0 / 72, 1.0 / 72};
  static constexpr int REV[15] = {0, 2, 1, 4, 3, 6, 5, 14, 13, 12, 11, 10, 9, 8, 7};
};

struct D3Q19 {
  static constexpr int Q = 19;
  static constexpr int E[19][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                                   {0, 0, 1},  {0, 0, -1},  {1, 1, 0},  {-1, 1, 0}, {1, -1, 0},
                                   {-1, -1, 0}, {0, 1, 1},  {0, -1, 1}, {0, 1, -1}, {0, -1, -1},
                                   {1, 0, 1},  {-1, 0, 1},  {1, 0, -1}, {-1, 0, -1}};
  static constexpr double W[19] = {1.0 / 3,  1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                                   1.0 / 18, 1.0 / 18, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                   1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                   1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
  static constexpr int REV[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
};

struct D3Q27 {
  static constexpr int Q = 27;
  static constexpr int E[27][3] = {
      {0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},  {0, 0, 1},  {0, 0, -1},
      {1, 1, 0},   {-1, 1, 0},  {1, -1, 0}, {-1, -1, 0}, {0, 1, 1},   {0, -1, 1}, {0, 1, -1},
      {0, -1, -1}, {1, 0, 1},   {-1, 0, 1}, {1, 0, -1},  {-1, 0, -1}, {1, 1, 1},  {1, 1, -1},
      {1, -1, 1},  {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
  static constexpr double W[27] = {
      8.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,  2.0 / 27,
      1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,
      1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 54,  1.0 / 216, 1.0 / 216,
      1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216, 1.0 / 216};
  static constexpr int REV[27] = {0,  2,  1,  4,  3,  6,  5,  10, 9,  8,  7,  14, 13, 12,
                                  11, 18, 17, 16, 15, 26, 25, 24, 23, 22, 21, 20, 19};
};

// ---------------------------------------------------------------------------
// Storage formats of the population state.
//   T      arithmetic type
//   elem   stored element
//   load   stored -> T for direction weight w
//   store  T -> stored
// Deviation formats keep (f - w*rho0)*scale; loads and stores are a
// multiply-add away from f, so streaming a deviation buffer is still a pure
// permutation of stored bits.
// ---------------------------------------------------------------------------
struct StoreF32 {
  using T = float;
  using elem = float;
  static constexpr int code = FUNLBM_STORE_F32;
  static constexpr bool fast = false;  // the reference's float32 operations, bit for bit
  static constexpr bool dev = false;   // loads give f (not a scaled deviation)
  using raw = float;
  __device__ static inline float load(elem v, float, float, float) { return v; }
  __device__ static inline elem store(float f, float, float, float) { return f; }
  __device__ static inline float load_raw(raw v, float, float, float) { return v; }
  __device__ static inline raw store_raw(float f, float, float, float) { return f; }
};

struct StoreF64 {
  using T = double;
  using elem = double;
  static constexpr int code = FUNLBM_STORE_F64;
  static constexpr bool fast = false;
  static constexpr bool dev = false;
  using raw = double;
  __device__ static inline double load(elem v, double, double, double) { return v; }
  __device__ static inline elem store(double f, double, double, double) { return f; }
  __device__ static inline double load_raw(raw v, double, double, double) { return v; }
  __device__ static inline raw store_raw(double f, double, double, double) { return f; }
};

struct StoreF16 {
  using T = float;
  using elem = __half;
  static constexpr int code = FUNLBM_STORE_F16;
  // 11-bit storage: its rounding (2^-11 of the deviation) dominates the
  // error, so the arithmetic uses reciprocals and FMAs (FastMath below)
  static constexpr bool fast = true;
  static constexpr bool dev = false;
  __device__ static inline float to_float(elem v) { return __half2float(v); }
  __device__ static inline unsigned short from_float(float g) { return __half_as_ushort(__float2half_rn(g)); }
  __device__ static inline float load(elem v, float wrho0, float inv_scale, float) {
    return __builtin_fmaf(__half2float(v), inv_scale, wrho0);
  }
  __device__ static inline elem store(float f, float wrho0, float, float scale) {
    return __float2half_rn((f - wrho0) * scale);
  }
  // stochastic rounding into the low (HI false) or high half of `old`: up by
  // one ulp iff the discarded bits, truncated to 13, plus rnd >> 19 carry out
  // of 13 bits (v_cvt_sr_f16_f32)
  template <bool HI>
  __device__ static inline uint32_t cvt_sr_half(uint32_t old, float g, uint32_t rnd) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(h2, old), g, rnd, HI));
  }
  // the same conversions on the 16-bit patterns (vector loads and stores)
  using raw = unsigned short;
  __device__ static inline float load_raw(raw v, float wrho0, float inv_scale, float s) {
    return load(__ushort_as_half(v), wrho0, inv_scale, s);
  }
  __device__ static inline raw store_raw(float f, float wrho0, float is, float scale) {
    return __half_as_ushort(store(f, wrho0, is, scale));
  }
};

struct StoreBF16 {
  using T = float;
  using elem = __hip_bfloat16;
  static constexpr int code = FUNLBM_STORE_BF16;
  static constexpr bool fast = true;
  static constexpr bool dev = false;
  __device__ static inline float to_float(elem v) { return __bfloat162float(v); }
  __device__ static inline unsigned short from_float(float g) {
    return __builtin_bit_cast(unsigned short, __float2bfloat16(g));
  }
  __device__ static inline float load(elem v, float wrho0, float inv_scale, float) {
    return __builtin_fmaf(__bfloat162float(v), inv_scale, wrho0);
  }
  __device__ static inline elem store(float f, float wrho0, float, float scale) {
    return __float2bfloat16(((f - wrho0) * scale));
  }
  // stochastic rounding into the low (HI false) or high half of `old`: up by
  // one ulp iff the 16 discarded bits plus rnd >> 16 carry out of 16 bits
  // (v_cvt_sr_bf16_f32)
  template <bool HI>
  __device__ static inline uint32_t cvt_sr_half(uint32_t old, float g, uint32_t rnd) {
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(b2, old), g, rnd, HI));
  }
  using raw = unsigned short;
  __device__ static inline float load_raw(raw v, float wrho0, float inv_scale, float s) {
    return load(__builtin_bit_cast(elem, v), wrho0, inv_scale, s);
  }
  __device__ static inline raw store_raw(float f, float wrho0, float is, float scale) {
    return __builtin_bit_cast(raw, store(f, wrho0, is, scale));
  }
};

// A 16-bit state read as its stored, scaled deviations g = (f - w_k) scale,
// without the conversion to f: the fused step's deviation arithmetic
// (moments_dev, collide_dev, add_forcing_dev) works on g directly.
template <class S>
struct DevOf {
  using T = float;
  using elem = typename S::elem;
  using raw = typename S::raw;
  static constexpr int code = S::code;
  static constexpr bool fast = true;
  static constexpr bool dev = true;
  __device__ static inline float load(elem v, float, float, float) { return S::to_float(v); }
  __device__ static inline raw store_raw(float g, float, float, float) { return S::from_float(g); }
};

// ---------------------------------------------------------------------------
// Dither of the stochastic rounding.  The 32-bit word of population k at node
// j in a step of seed s is the low 32 bits of (h mod 2^24) * C_k with
// h = mix32(j ^ mix32(s)) (one v_mul_u32_u24 per population).  The
// conversions use the word's top 13 (f16) or 16 (bf16) bits: a multiplicative
// hash of h per direction.  C_k are odd multipliers in [2^23, 2^24), taken in
// order from (mix32(i) mod 2^24) | 0x800001, i = 1, 2, ..., skipping any whose
// top bits correlate with an earlier one's by more than 0.012 over 3 x 2^16
// nodes (multipliers in a near small-integer ratio, such as C' = 2C - 1, give
// correlated dithers).
// mix32 is the 32-bit integer hash x ^= x >> 16; x *= 0x7feb352d;
// x ^= x >> 15; x *= 0x846ca68b; x ^= x >> 16 (C. Wellons' "lowbias32").
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// The seed of the step after `steps_done` steps from base seed `base`.
__host__ __device__ inline uint32_t step_seed(uint32_t base, uint32_t steps_done) {
  return base + steps_done * 0x9E3779B9u;
}

__host__ __device__ constexpr uint32_t dither_mult(int k) {
  constexpr uint32_t C[27] = {0x8990c1, 0x932181, 0xf1e9dd, 0xfe5ed1, 0xc5d53f, 0xe3d3bb, 0x8ba1e7,
                              0xd35fbb, 0x9fd5ab, 0x8baa7d, 0xa8f9fb, 0xdc746d, 0xbfcff1, 0xf63447,
                              0xbd4a6f, 0xa56ae5, 0xabdd59, 0xbb2f91, 0x8001ed, 0xdd7671, 0xbe2e8f,
                              0xdbcac3, 0x9006fb, 0xce98a9, 0xff9fe3, 0xbc09cb, 0xdc5b09};
  return C[k];
}

// ---------------------------------------------------------------------------
// Kernel-side lattice parameters, built from funlbm_lattice on the host.
// ---------------------------------------------------------------------------
struct KParams {
  int nx, ny, nz;
  int sxy;    // ny * nz, the x stride of the node index
  int64_t n;  // nodes (< 2^31, checked on the host)
  int ts;     // tile shift of the fused-state layout (state_index)
  // Per direction k, the pull source node - s_k (s_k = e_k . (sxy, nz, 1))
  // split as s_k = q_k T + r_k, 0 <= r_k < T: soff[k] = byte offset of element
  // (k, tile -q_k, in-tile -r_k) from the state's tile 0 minus the slack, i.e.
  // ((Q - q_k Q + k) T - r_k) * element size (|soff| < 2^31, checked on the
  // host).  r_k * element size = (-soff[k]) mod (T * element size).
  int32_t soff[27];
  // bit w: the sources of some direction straddle a tile boundary for the
  // wave at in-tile start 64 w (some r_k in (64 w, 64 w + 64))
  uint64_t mixmask;
  // node -> (x, y, z) by multiply-shift: q = (n * m) >> s is floor(n / d) for
  // every n < 2^31 (Granlund-Montgomery; m = floor(2^s / d) + 1, s = 31 + ceil(log2 d))
  uint32_t mz, my;
  int sz, sy;
  int bc[6];
  int uzero;
  int has_moving;  // any WALL_WITH_SPEED face
  int has_neq;     // a NON_EQUILIBRIUM input/output face (explicit path only)
  int has_body;    // body_force != 0
  int sr;          // stochastic rounding of a 16-bit state's stores
  int delta;       // FUNLBM_DELTA_*: the immersed-boundary kernel
  uint32_t seed;   // the lattice's dither seed (the base, for a batched case)
  uint32_t seed_mix;  // dither_mix(seed): this step's mix (single-lattice calls)
  float tau;       // float32 tau as the reference's field
  float rtau;      // float32(1 / tau)
  float omega;     // 1 - 1/(2 tau) in float32, the forcing prefactor
  double tau_d, rtau_d, omega_d;  // the same in double for float64 storage
  float rho0, scale, inv_scale;
  float g[3];
  float uw[6][3];
};

// Relaxation constants in the arithmetic type T.
template <class T>
struct Relax {
  T tau, rtau, omega;
};
template <class T>
__device__ inline Relax<T> relax_of(const KParams& p);
template <>
__device__ inline Relax<float> relax_of<float>(const KParams& p) {
  return {p.tau, p.rtau, p.omega};
}
template <>
__device__ inline Relax<double> relax_of<double>(const KParams& p) {
  return {p.tau_d, p.rtau_d, p.omega_d};
}

// x / c correctly rounded in three operations, given rc = RN(1/c): q = x rc,
// the residual x - q c is exact with an FMA, and q + residual * rc rounds to
// RN(x / c) (Markstein's theorem; binary floating point, round to nearest,
// results away from the subnormal range).  Used wherever the reference
// divides, so quotients stay bit-identical to a true division.
__device__ __forceinline__ float divc(float x, float c, float rc) {
  const float q = x * rc;
  const float e = __builtin_fmaf(-q, c, x);
  return __builtin_fmaf(e, rc, q);
}
__device__ __forceinline__ double divc(double x, double c, double rc) {
  const double q = x * rc;
  const double e = __builtin_fma(-q, c, x);
  return __builtin_fma(e, rc, q);
}

// Constants of the reference: cs = float32(sqrt(1/3)), cs^2 and cs^4 rounded
// to float32 as torch computes `cs**2` and `cs2**2` on float32 scalars
// (src/funlbm/flow/base.py:37-39, flow/d3/base.py:106-107,152-156).
template <class T>
struct Cs;
template <>
struct Cs<float> {
  __device__ static inline float cs2() { return __uint_as_float(0x3eaaaaaau); }   // fl32(cs*cs)
  __device__ static inline float cs4() { return __uint_as_float(0x3de38e37u); }   // fl32(cs2*cs2)
  __device__ static inline float rcs2() { return __uint_as_float(0x40400001u); }  // RN(1/cs2)
  __device__ static inline float rcs4() { return __uint_as_float(0x41100001u); }  // RN(1/cs4)
};
template <>
struct Cs<double> {
  __device__ static inline double cs2() { return 1.0 / 3.0; }
  __device__ static inline double cs4() { return 1.0 / 9.0; }
  __device__ static inline double rcs2() { return 3.0; }  // RN(1/(1/3)) in double
  __device__ static inline double rcs4() { return 9.0; }  // RN(1/(1/9)) in double
};

template <class T>
__device__ inline T wt(double w) {
  return (T)w;
}

// Node coordinates from a flattened node index (z fastest).
struct Node {
  int x, y, z;
};
__device__ inline uint32_t fastdiv(uint32_t n, uint32_t m, int s) {
  return (uint32_t)(((uint64_t)n * m) >> s);
}

__device__ inline Node node_coords(const KParams& p, int64_t node) {
  Node c;
  const uint32_t nd = (uint32_t)node;
  const uint32_t t = fastdiv(nd, p.mz, p.sz);
  c.z = (int)(nd - t * (uint32_t)p.nz);
  const uint32_t x = fastdiv(t, p.my, p.sy);
  c.y = (int)(t - x * (uint32_t)p.ny);
  c.x = (int)x;
  return c;
}

__device__ inline bool interior_node(const KParams& p, const Node& c) {
  return (c.x > 0) & (c.x < p.nx - 1) & (c.y > 0) & (c.y < p.ny - 1) & (c.z > 0) & (c.z < p.nz - 1);
}

__device__ inline bool on_face(const KParams& p, const Node& c, int face) {
  switch (face) {
    case 0: return c.x == 0;
    case 1: return c.x == p.nx - 1;
    case 2: return c.y == 0;
    case 3: return c.y == p.ny - 1;
    case 4: return c.z == 0;
    default: return c.z == p.nz - 1;
  }
}

__device__ inline bool uzero_at(const KParams& p, const Node& c) {
  if (!p.uzero) return false;
  bool z = false;
#pragma unroll
  for (int f = 0; f < 6; ++f) z |= ((p.uzero >> f) & 1) && on_face(p, c, f);
  return z;
}

// ---------------------------------------------------------------------------
// Source of population k at node c after streaming.  The reference streams by
// torch.roll (push: f[x] = f*[x - e_k], periodic wrap) and then overrides the
// six faces in the order input, output, back, forward, bottom, top; a later
// override wins (src/funlbm/flow/d3/base.py:161-323).  Per face and its
// incoming directions (e_axis = +1 on the low face, -1 on the high face):
//   PERIODICAL       f*[opposite face, same transverse coordinates, k]
//   WALL             f*[same node, rev(k)]
//   WALL_WITH_SPEED  f*[same node, rev(k)] - 2 rho w_k (e_k . uw)/cs^2
// ---------------------------------------------------------------------------
struct Src {
  int x, y, z, k;
  int moving_face;  // -1 or the face whose wall velocity corrects the value
  int neq;          // 1 / -1: the input / output NON_EQUILIBRIUM rule gives the value
};

template <class L, int K>
__device__ inline Src source_of(const KParams& p, const Node& c) {
  constexpr int ex = L::E[K][0], ey = L::E[K][1], ez = L::E[K][2];
  Src s;
  s.k = K;
  s.moving_face = -1;
  s.neq = 0;
  s.x = c.x - ex;
  s.y = c.y - ey;
  s.z = c.z - ez;
  if (ex == 1 && c.x == 0) s.x = p.nx - 1;
  if (ex == -1 && c.x == p.nx - 1) s.x = 0;
  if (ey == 1 && c.y == 0) s.y = p.ny - 1;
  if (ey == -1 && c.y == p.ny - 1) s.y = 0;
  if (ez == 1 && c.z == 0) s.z = p.nz - 1;
  if (ez == -1 && c.z == p.nz - 1) s.z = 0;

  auto apply = [&](int face, int ox, int oy, int oz) {
    const int bc = p.bc[face];
    if (bc == FUNLBM_BC_NON_EQUILIBRIUM) return;  // set for every direction below
    s.neq = 0;
    if (bc == FUNLBM_BC_PERIODICAL) {
      s.x = ox; s.y = oy; s.z = oz; s.k = K; s.moving_face = -1;
    } else {  // WALL or WALL_WITH_SPEED (others are rejected on the host)
      s.x = c.x; s.y = c.y; s.z = c.z; s.k = L::REV[K];
      s.moving_face = (bc == FUNLBM_BC_WALL_WITH_SPEED) ? face : -1;
    }
  };
  // NON_EQUILIBRIUM input/output: every direction of the face plane, in the
  // reference's order (input, then output), before the y and z faces
  if (c.x == 0 && p.bc[0] == FUNLBM_BC_NON_EQUILIBRIUM) s.neq = 1;
  if (ex == 1 && c.x == 0) apply(0, p.nx - 1, c.y, c.z);
  if (ex == -1 && c.x == p.nx - 1) apply(1, 0, c.y, c.z);
  if (c.x == p.nx - 1 && p.bc[1] == FUNLBM_BC_NON_EQUILIBRIUM) s.neq = -1;
  if (ey == 1 && c.y == 0) apply(2, c.x, p.ny - 1, c.z);
  if (ey == -1 && c.y == p.ny - 1) apply(3, c.x, 0, c.z);
  if (ez == 1 && c.z == 0) apply(4, c.x, c.y, p.nz - 1);
  if (ez == -1 && c.z == p.nz - 1) apply(5, c.x, c.y, 0);
  return s;
}

__device__ inline int64_t lin(const KParams& p, int x, int y, int z) {
  return (int64_t)(x * p.sxy + y * p.nz + z);
}

// Fused-state layout.  Nodes are grouped into tiles of T = 2^ts consecutive
// nodes; a tile holds its nodes' Q direction planes one after another (T
// elements each).  One slack tile precedes tile 0 and one follows the last, so
// the one-node overreach of a corner direction's shifted load on a z-face node
// (node -1 or node n, whose value the face rule then replaces) stays inside the
// buffer.  Element (k, j) for j in [-1, n]:
//   Q T + ((j >> ts) Q + k) T + (j & (T - 1)).
template <int Q>
__device__ inline int64_t state_index(const KParams& p, int k, int64_t j) {
  const int64_t tile = j >> p.ts;  // floor (arithmetic shift), -1 for j = -1
  return ((int64_t)Q + tile * Q + k) * ((int64_t)1 << p.ts) + (j & (((int64_t)1 << p.ts) - 1));
}

// Moving-wall correction 2 (rho w_k) (e_k . uw) / cs^2 in the reference's
// operation order (src/funlbm/flow/d3/base.py:204-213).
template <class L, int K, class T>
__device__ inline T moving_wall_term(const KParams& p, int face, T rho) {
  constexpr int ex = L::E[K][0], ey = L::E[K][1], ez = L::E[K][2];
  T eu = (T)0;
  eu = eu + (T)ex * (T)p.uw[face][0];
  eu = eu + (T)ey * (T)p.uw[face][1];
  eu = eu + (T)ez * (T)p.uw[face][2];
  T tmp = (T)2 * (rho * wt<T>(L::W[K]));
  return tmp * eu / Cs<T>::cs2();
}

// The z-face override of a node whose x and y are inner (only the bottom or
// top rule of source_of can apply there): directions with e_z = DZ take, on a
// periodic face, plane K at the same (x, y) on the opposite z face; on a wall,
// plane rev(K) at the node itself, less the moving-wall term.  The lanes on
// the face issue the loads of all those directions before the first use, so
// a wave waits for one round trip, not one per direction; the values then go
// into f by a select, outside the guarded block (a store into f under the
// guard lets the compiler merge the two faces' stores through a pointer
// select, which puts the array in private memory).
template <class L, int DZ>
struct ZDirs {
  static constexpr int count() {
    int n = 0;
    for (int k = 0; k < L::Q; ++k) n += L::E[k][2] == DZ;
    return n;
  }
  // the i-th direction with e_z = DZ
  static constexpr int dir(int i) {
    for (int k = 0; k < L::Q; ++k)
      if (L::E[k][2] == DZ && i-- == 0) return k;
    return -1;
  }
};

template <class L, class S, int DZ>
__device__ inline void zface_pull(const KParams& p, const typename S::elem* __restrict__ src, int64_t node,
                                  const float* __restrict__ rho_prev, bool hit, typename S::T* f) {
  using T = typename S::T;
  using E = typename S::elem;
  using Z = ZDirs<L, DZ>;
  constexpr int NZ = Z::count();
  constexpr int face = DZ > 0 ? 4 : 5;
  const int bc = p.bc[face];
  T rep[NZ];
  if (hit) {
    E raw[NZ];
    if (bc == FUNLBM_BC_PERIODICAL) {
      const int64_t across = node + (DZ > 0 ? (p.nz - 1) : -(p.nz - 1));
      static_for<0, NZ>([&](auto ic) {
        constexpr int I = decltype(ic)::value;
        raw[I] = src[state_index<L::Q>(p, Z::dir(I), across)];
      });
    } else {
      static_for<0, NZ>([&](auto ic) {
        constexpr int I = decltype(ic)::value;
        raw[I] = src[state_index<L::Q>(p, L::REV[Z::dir(I)], node)];
      });
    }
    static_for<0, NZ>([&](auto ic) {
      constexpr int I = decltype(ic)::value;
      constexpr int K = Z::dir(I);
      rep[I] = S::load(raw[I], wt<float>(L::W[K]), p.inv_scale, p.scale);
    });
    if (bc == FUNLBM_BC_WALL_WITH_SPEED) {
      const T rho = (T)rho_prev[node];
      static_for<0, NZ>([&](auto ic) {
        constexpr int I = decltype(ic)::value;
        constexpr int K = Z::dir(I);
        if constexpr (S::dev)
          rep[I] = rep[I] - moving_wall_term<L, K, T>(p, face, rho) * p.scale;
        else
          rep[I] = rep[I] - moving_wall_term<L, K, T>(p, face, rho);
      });
    }
  }
  static_for<0, NZ>([&](auto ic) {
    constexpr int I = decltype(ic)::value;
    constexpr int K = Z::dir(I);
    f[K] = hit ? rep[I] : f[K];
  });
}

// Pull all Q post-stream populations of node c out of a state buffer.
// Nodes with inner x and y load every plane at the shifted source node - e_k
// (in bounds even on a z face, where the row wraps into the neighbouring
// row), then a z-face node replaces its e_z-pointing directions by the face
// rule; nodes on an x or y face go through the full override logic of
// source_of.  In node order a wave covers part of one z-row, so the x/y-face
// rows are whole waves and only the z-face lanes take the short fix-up.
//
// ALIGNED: the 64 lanes of the wave hold 64 consecutive nodes starting at a
// multiple of 64 (every full-lattice kernel).  The wave then lies in one
// state tile, and the source of direction K, node - s_K, lies in one tile or
// straddles two: the tile and in-tile start are wave-uniform scalars and a
// lane adds only its in-tile byte offset, as in a plain copy.
template <class L, class S, bool ALIGNED = true>
__device__ inline void pull_node(const KParams& p, const typename S::elem* __restrict__ src,
                                 const Node& c, int64_t node, const float* __restrict__ rho_prev,
                                 typename S::T* f) {
  using T = typename S::T;
  using E = typename S::elem;
  constexpr int Q = L::Q;
  if ((c.x > 0) & (c.x < p.nx - 1) & (c.y > 0) & (c.y < p.ny - 1)) {
    if constexpr (ALIGNED) {
      const int ts = p.ts;
      const uint32_t tmask = (1u << ts) - 1u;
      const int nw = __builtin_amdgcn_readfirstlane((int)node) & ~63;  // the wave's first node
      const uint32_t l0 = (uint32_t)nw & tmask;                        // its in-tile start
      const uint32_t local = (uint32_t)node & tmask;
      const uint32_t lb = local * (uint32_t)sizeof(E);
      // the wave's tile, and one tile of Q planes less one plane (tile t,
      // plane K -> tile t - 1, plane K)
      const char* tb = (const char*)src + (int64_t)(nw >> ts) * Q * ((int64_t)sizeof(E) << ts);
      const uint32_t back = (uint32_t)((Q - 1) * ((int64_t)sizeof(E) << ts));
      const uint32_t tbmask = ((uint32_t)sizeof(E) << ts) - 1u;  // bytes of one plane of a tile, less one
      const uint32_t l0b = l0 * (uint32_t)sizeof(E);
      if (!((p.mixmask >> (l0 >> 6)) & 1)) {
        // every direction's sources lie in one tile: scalar base + the lane's
        // in-tile offset, tile tn - q_K (r_K <= l0) or tn - q_K - 1
        static_for<0, Q>([&](auto kc) {
          constexpr int K = decltype(kc)::value;
          const int32_t so = p.soff[K];
          const uint32_t rkb = (uint32_t)(-so) & tbmask;  // r_K in bytes
          const char* ua = tb + so;  // element (K, tile tn - q_K, in-tile -r_K)
          const char* sb = rkb <= l0b ? ua : ua - back;
          f[K] = S::load(*(const E*)(sb + lb), wt<float>(L::W[K]), p.inv_scale, p.scale);
        });
      } else {
        // a wave at a tile edge: per lane, the source tile of a straddling direction
        static_for<0, Q>([&](auto kc) {
          constexpr int K = decltype(kc)::value;
          const int32_t so = p.soff[K];
          const uint32_t rkb = (uint32_t)(-so) & tbmask;
          const char* ua = tb + so;
          const char* sb = rkb <= l0b ? ua : ua - back;
          const uint32_t vo = ((rkb > l0b) & (rkb < l0b + 64u * (uint32_t)sizeof(E))) ? lb + (lb >= rkb ? back : 0u) : lb;
          f[K] = S::load(*(const E*)(sb + vo), wt<float>(L::W[K]), p.inv_scale, p.scale);
        });
        // keeps the compiler from sinking both paths' loads into one block
        // behind per-lane 64-bit addresses: the common path above then loads
        // from a scalar base plus the lane's 32-bit offset
        __asm__ volatile("" ::: "memory");
      }
    } else {
      static_for<0, Q>([&](auto kc) {
        constexpr int K = decltype(kc)::value;
        constexpr int ex = L::E[K][0], ey = L::E[K][1], ez = L::E[K][2];
        const int64_t sn = node - (ex * p.sxy + ey * p.nz + ez);
        f[K] = S::load(src[state_index<Q>(p, K, sn)], wt<float>(L::W[K]), p.inv_scale,
                       p.scale);
      });
    }
    zface_pull<L, S, 1>(p, src, node, rho_prev, c.z == 0, f);
    zface_pull<L, S, -1>(p, src, node, rho_prev, c.z == p.nz - 1, f);
    return;
  }
  // every direction's load is issued before the first is used (one round
  // trip per node, not one per direction); the moving-wall terms follow
  E raw[Q];
  static_for<0, Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    const Src s = source_of<L, K>(p, c);
    raw[K] = src[state_index<Q>(p, s.k, lin(p, s.x, s.y, s.z))];
  });
  // The element was stored with the weight of s.k; rev(k) and k share a
  // weight on every lattice.
  static_for<0, Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    f[K] = S::load(raw[K], wt<float>(L::W[K]), p.inv_scale, p.scale);
  });
  if (p.has_moving) {
    const T rho = (T)rho_prev[node];
    static_for<0, Q>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      const Src s = source_of<L, K>(p, c);
      if (s.moving_face >= 0) {
        if constexpr (S::dev)
          f[K] = f[K] - moving_wall_term<L, K, T>(p, s.moving_face, rho) * p.scale;
        else
          f[K] = f[K] - moving_wall_term<L, K, T>(p, s.moving_face, rho);
      }
    });
  }
}

// ---------------------------------------------------------------------------
// Moments (src/funlbm/flow/d3/base.py:116-133): rho = sum_k f_k in direction
// order; j = f.e in direction order; u = (j + F/2)/rho; u := 0 on masked faces.
// ---------------------------------------------------------------------------
template <class L, class T, bool FAST = false>
__device__ inline void moments(const KParams& p, const Node& c, const T* f, const T F[3], T& rho,
                               T u[3]) {
  T r = (T)0, jx = (T)0, jy = (T)0, jz = (T)0;
#pragma unroll
  for (int k = 0; k < L::Q; ++k) {
    r = r + f[k];
    if (L::E[k][0] == 1) jx = jx + f[k];
    if (L::E[k][0] == -1) jx = jx - f[k];
    if (L::E[k][1] == 1) jy = jy + f[k];
    if (L::E[k][1] == -1) jy = jy - f[k];
    if (L::E[k][2] == 1) jz = jz + f[k];
    if (L::E[k][2] == -1) jz = jz - f[k];
  }
  rho = r;
  if constexpr (FAST) {
    // the hardware reciprocal (1 ulp) and products
    const T rr = __builtin_amdgcn_rcpf(r);
    u[0] = __builtin_fmaf(F[0], (T)0.5, jx) * rr;
    u[1] = __builtin_fmaf(F[1], (T)0.5, jy) * rr;
    u[2] = __builtin_fmaf(F[2], (T)0.5, jz) * rr;
  } else {
    const T rr = (T)1 / r;
    u[0] = divc(jx + F[0] / (T)2, r, rr);
    u[1] = divc(jy + F[1] / (T)2, r, rr);
    u[2] = divc(jz + F[2] / (T)2, r, rr);
  }
  if (uzero_at(p, c)) {
    u[0] = (T)0;
    u[1] = (T)0;
    u[2] = (T)0;
  }
}

// e_k . u in the reference's summation order (x, y, z), exact up to the two
// additions; tmp_k = (e_k . u)/cs^2.
template <class L, int K, class T>
__device__ inline T edotu(const T u[3]) {
  T d = (T)0;
  if (L::E[K][0] == 1) d = d + u[0];
  if (L::E[K][0] == -1) d = d - u[0];
  if (L::E[K][1] == 1) d = d + u[1];
  if (L::E[K][1] == -1) d = d - u[1];
  if (L::E[K][2] == 1) d = d + u[2];
  if (L::E[K][2] == -1) d = d - u[2];
  return d;
}

// The first direction with a nonzero component along axis d.
template <class L>
__host__ __device__ constexpr int first_along(int d) {
  for (int k = 0; k < L::Q; ++k)
    if (L::E[k][d] != 0) return k;
  return 0;
}

// e_k . u without the leading zero of edotu (the deviation path, where the
// sign of a zero product does not matter): one add or subtract fewer.
template <class L, int K>
__device__ inline float edotu_dev(const float u[3]) {
  constexpr int a = L::E[K][0], b = L::E[K][1], c = L::E[K][2];
  float d = a == 1 ? u[0] : a == -1 ? -u[0] : b == 1 ? u[1] : b == -1 ? -u[1] : c == 1 ? u[2] : -u[2];
  if constexpr (a != 0) {
    if constexpr (b == 1) d = d + u[1];
    if constexpr (b == -1) d = d - u[1];
  }
  if constexpr (a != 0 || b != 0) {
    if constexpr (c == 1) d = d + u[2];
    if constexpr (c == -1) d = d - u[2];
  }
  return d;
}

template <class T>
__device__ inline T sqrt_t(T v);
// Correctly rounded square roots, as numpy and torch compute them.  The
// compiler's lowering of the sqrt builtin refines the hardware's faithful
// v_sqrt_f32 with the residual test of its two neighbours; ROCm's __fsqrt_rn
// is the bare faithful instruction and differs in the last bit now and then.
template <>
__device__ inline float sqrt_t<float>(float v) {
  return __builtin_sqrtf(v);
}
template <>
__device__ inline double sqrt_t<double>(double v) {
  return __builtin_sqrt(v);
}

// BGK collision (src/funlbm/flow/d3/base.py:150-159).  feq_k =
// (1 + tmp + tmp^2/2 - u2) * (rho w_k), tmp = (u.e_k)/cs^2,
// u2 = |u|^2/cs^2/2 with |u| = sqrt(u.u) squared again as torch does.
// f_k <- f_k - (f_k - feq_k)/tau.  If feq != nullptr it receives feq_k.
// e_rev(k) . u is exactly -(e_k . u), so tmp of a reverse direction is the
// negated tmp of its partner.
template <class L, class T, bool FAST = false>
__device__ inline void collide(T* f, const T u[3], T rho, const Relax<T>& rx, T* feq) {
  if constexpr (FAST) {
    // the same equilibrium with 1/cs^2 = 3 and 1/cs^4 = 9 as products,
    // |u|^2 without the square root, and the relaxation as one FMA:
    // feq_k = (1 + t (1 + t/2) - 3/2 |u|^2) rho w_k, f_k += (feq_k - f_k)/tau
    const T uu = __builtin_fmaf(u[2], u[2], __builtin_fmaf(u[1], u[1], u[0] * u[0]));
    const T u2 = uu * (T)1.5;
    T tmp[L::Q];
    static_for<0, L::Q>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      if constexpr (L::REV[K] < K)
        tmp[K] = -tmp[L::REV[K]];
      else
        tmp[K] = edotu<L, K, T>(u) * (T)3;
    });
    static_for<0, L::Q>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      const T t = tmp[K];
      const T e = (__builtin_fmaf(t, __builtin_fmaf(t, (T)0.5, (T)1), (T)1) - u2) * (rho * wt<T>(L::W[K]));
      if (feq) feq[K] = e;
      f[K] = __builtin_fmaf(e - f[K], rx.rtau, f[K]);
    });
    return;
  }
  const T nrm = sqrt_t<T>(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  const T u2 = divc(nrm * nrm, Cs<T>::cs2(), Cs<T>::rcs2()) / (T)2;
  T tmp[L::Q];
  static_for<0, L::Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    if constexpr (L::REV[K] < K)
      tmp[K] = -tmp[L::REV[K]];
    else
      tmp[K] = divc(edotu<L, K, T>(u), Cs<T>::cs2(), Cs<T>::rcs2());
  });
  static_for<0, L::Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    const T t = tmp[K];
    const T e = ((((T)1 + t) + t * t / (T)2) - u2) * (rho * wt<T>(L::W[K]));
    if (feq) feq[K] = e;
    f[K] = f[K] - divc(f[K] - e, rx.tau, rx.rtau);
  });
}

// Forcing term of FlowD3.cul_equ2 (src/funlbm/flow/d3/base.py:103-114):
// t1 = (e_a - u)/cs^2, t2 = (u e_a) e_a / cs^4 (elementwise),
// t3 = sum_d (t1 + t2)_d F_d, f_a += ((1 - 1/(2 tau)) w_a) t3.
template <class L, class T, bool FAST = false>
__device__ inline void add_forcing(T* f, const T u[3], const T F[3], const Relax<T>& rx) {
  const T omega = rx.omega;
  if constexpr (FAST) {
    // (t1 + t2)_d = 3 (e_d - u_d) + 9 u_d e_d^2: 3 e_d + 6 u_d where e_d != 0,
    // -3 u_d where e_d = 0; so t3 = sum_d of A_d e_d + B_d or C_d
    T A[3], B[3], C[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      A[d] = (T)3 * F[d];
      B[d] = ((T)6 * u[d]) * F[d];
      C[d] = ((T)-3 * u[d]) * F[d];
    }
    static_for<0, L::Q>([&](auto kc) {
      constexpr int K = decltype(kc)::value;
      T t3 = (T)0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int e = L::E[K][d];
        t3 = t3 + (e == 0 ? C[d] : (e > 0 ? A[d] + B[d] : B[d] - A[d]));
      }
      f[K] = __builtin_fmaf(omega * wt<T>(L::W[K]), t3, f[K]);
    });
    return;
  }
  static_for<0, L::Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    T t3 = (T)0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T e = (T)L::E[K][d];
      const T t1 = divc(e - u[d], Cs<T>::cs2(), Cs<T>::rcs2());
      const T t2 = divc((u[d] * e) * e, Cs<T>::cs4(), Cs<T>::rcs4());
      t3 = (d == 0) ? (t1 + t2) * F[d] : t3 + (t1 + t2) * F[d];
    }
    f[K] = f[K] + (omega * wt<T>(L::W[K])) * t3;
  });
}

// ---------------------------------------------------------------------------
// Deviation arithmetic of the 16-bit states: g_k = (f_k - w_k) s with s the
// storage scale (rho0 = 1).  Since sum_k w_k = 1 and sum_k w_k e_k = 0:
// rho = 1 + sum_k g_k / s, j = sum_k e_k g_k / s; the equilibrium deviation is
// g_eq,k = s w_k (rho A_k - 1) with A_k = 1 + t_k (1 + t_k / 2) - 3/2 |u|^2,
// t_k = 3 e_k . u; the forcing adds s omega w_k t3 (add_forcing, FAST).  The
// same operations as the FAST path on f, without converting each population
// to f and back (38 fewer instructions per node), and on smaller numbers.
// ---------------------------------------------------------------------------
// drho (optional): rho - 1 from the sum of the deviations, to float32's
// relative precision; rho, held as 1 + drho in float32, resolves drho only to
// about 6e-8.
template <class L>
__device__ inline void moments_dev(const KParams& p, const Node& c, const float* g, const float F[3], float& rho,
                                   float u[3], float* drho = nullptr) {
  // each sum starts from its first term (no leading zero to add)
  constexpr int fx = first_along<L>(0), fy = first_along<L>(1), fz = first_along<L>(2);
  float r = g[0];
  float jx = L::E[fx][0] == 1 ? g[fx] : -g[fx];
  float jy = L::E[fy][1] == 1 ? g[fy] : -g[fy];
  float jz = L::E[fz][2] == 1 ? g[fz] : -g[fz];
  static_for<1, L::Q>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    r = r + g[k];
    if constexpr (k != fx && L::E[k][0] == 1) jx = jx + g[k];
    if constexpr (k != fx && L::E[k][0] == -1) jx = jx - g[k];
    if constexpr (k != fy && L::E[k][1] == 1) jy = jy + g[k];
    if constexpr (k != fy && L::E[k][1] == -1) jy = jy - g[k];
    if constexpr (k != fz && L::E[k][2] == 1) jz = jz + g[k];
    if constexpr (k != fz && L::E[k][2] == -1) jz = jz - g[k];
  });
  rho = __builtin_fmaf(r, p.inv_scale, 1.0f);
  if (drho) *drho = r * p.inv_scale;
  const float rr = __builtin_amdgcn_rcpf(rho);
  u[0] = __builtin_fmaf(F[0], 0.5f, jx * p.inv_scale) * rr;
  u[1] = __builtin_fmaf(F[1], 0.5f, jy * p.inv_scale) * rr;
  u[2] = __builtin_fmaf(F[2], 0.5f, jz * p.inv_scale) * rr;
  if (uzero_at(p, c)) {
    u[0] = 0.0f;
    u[1] = 0.0f;
    u[2] = 0.0f;
  }
}

// g <- g (1 - 1/tau) + (s/tau) w_k (rho a_k - 1), the BGK relaxation towards
// g_eq = s w_k (rho a_k - 1), a_k = 1 + t + t^2/2 - 3|u|^2/2, t = 3 e_k.u.
// With c0 = rho (1 - 3|u|^2/2) - 1 = drho - (3/2)|u|^2 rho, rho a_k - 1 =
// c0 + rho t (1 + t/2).  c0 is formed from drho = rho - 1 (the sum of the
// deviations), not from rho: at slow flow its two terms lie far below
// float32's resolution of rho near 1.  A direction and its reverse
// (t -> -t) share, with sw = (s/tau) w_k, anti = sw rho t = (3 sw rho)(e_k.u)
// and sym = sw c0 + anti (3/2)(e_k.u), and take sym +- anti; 3 sw rho and
// sw c0 are formed once per weight.
template <class L>
__device__ inline void collide_dev(const KParams& p, float* g, const float u[3], float rho, float drho) {
  const float uu = __builtin_fmaf(u[2], u[2], __builtin_fmaf(u[1], u[1], u[0] * u[0]));
  const float c0 = __builtin_fmaf(uu * -1.5f, rho, drho);
  const float sr = p.scale * p.rtau;  // s / tau
  const float keep = 1.0f - p.rtau;
  static_for<0, L::Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    constexpr int R = L::REV[K];
    const float sw = sr * wt<float>(L::W[K]);
    if constexpr (R == K) {
      g[K] = __builtin_fmaf(g[K], keep, sw * c0);
    } else if constexpr (R > K) {
      const float eu = edotu_dev<L, K>(u);
      const float anti = ((3.0f * sw) * rho) * eu;
      const float sym = __builtin_fmaf(anti, 1.5f * eu, sw * c0);
      g[K] = __builtin_fmaf(g[K], keep, sym + anti);
      g[R] = __builtin_fmaf(g[R], keep, sym - anti);
    }
  });
}

template <class L>
__device__ inline void add_forcing_dev(const KParams& p, float* g, const float u[3], const float F[3]) {
  float A[3], B[3], C[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    A[d] = 3.0f * F[d];
    B[d] = (6.0f * u[d]) * F[d];
    C[d] = (-3.0f * u[d]) * F[d];
  }
  const float so = p.scale * p.omega;
  static_for<0, L::Q>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    float t3 = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int e = L::E[K][d];
      t3 = t3 + (e == 0 ? C[d] : (e > 0 ? A[d] + B[d] : B[d] - A[d]));
    }
    g[K] = __builtin_fmaf(so * wt<float>(L::W[K]), t3, g[K]);
  });
}

// Block index remap so that consecutive logical blocks share one XCD (and its
// L2): the hardware hands out workgroups round-robin over the 8 XCDs.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nblocks) {
  constexpr uint32_t NX = 8;
  const uint32_t per = nblocks / NX;
  if (b >= per * NX) return b;  // the ragged tail keeps its index
  return (b % NX) * per + b / NX;
}

}  // namespace funlbm
