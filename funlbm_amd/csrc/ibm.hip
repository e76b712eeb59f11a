// Immersed-boundary coupling to rigid particles: marker interpolation and
// direct forcing, wall spring, force/torque reduction and rigid-body update in
// double precision, pose update, and spreading with the reference's overwrite
// semantics made deterministic.
//
// Markers of all particles are concatenated (particle-major, the reference's
// loop order).  One 64-lane wavefront serves one marker: lane (i, j, k) is node
// rl + (i, j, k) of the 4 x 4 x 4 stencil.
#include "capi_util.h"
#include "lbm_common.h"

namespace funlbm {

constexpr int kWave = 64;
constexpr int kIbmBlock = 256;  // 4 markers per block
constexpr float kPiF = 3.14159265358979323846f;

// Stencil of a marker (src/funlbm/lbm/lbm3d.py:32-44): rl = max(floor(x) - 1, 0),
// rr = min(rl + 4, size) with size = the config size (node count - 1).
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
    float rr = fminf(rl + 4.0f, (float)size[d]);
    s.lo[d] = (int)rl;
    s.hi[d] = (int)rr;
  }
  return s;
}

// Cosine kernel factor (src/funlbm/lbm/lbm3d.py:46-59):
// phi(r) = (1 + cos(|r pi/2|))/4 where all |r_d| <= 2, else 0.
__device__ inline float unnormalised_weight(const float x[3], int nx_, int ny_, int nz_, bool& inside) {
  const float r[3] = {(float)nx_ - x[0], (float)ny_ - x[1], (float)nz_ - x[2]};
  inside = fabsf(r[0]) <= 2.0f && fabsf(r[1]) <= 2.0f && fabsf(r[2]) <= 2.0f;
  if (!inside) return 0.0f;
  float t[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) t[d] = (1.0f + cosf(fabsf(r[d] * kPiF / 2.0f))) / 4.0f;
  return (t[0] * t[1]) * t[2];
}

template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Lane geometry of marker m: node coordinates and normalised weight.
struct LaneW {
  bool valid;
  int x, y, z;
  float w;     // normalised weight (0 for invalid lanes)
  float wsum;  // sum of normalised weights of the stencil
};

__device__ inline LaneW lane_weight(const KParams& p, const float x[3], int lane) {
  LaneW lw;
  const Stencil s = stencil_of(p, x);
  const int i = lane >> 4, j = (lane >> 2) & 3, k = lane & 3;
  lw.x = s.lo[0] + i;
  lw.y = s.lo[1] + j;
  lw.z = s.lo[2] + k;
  lw.valid = lw.x < s.hi[0] && lw.y < s.hi[1] && lw.z < s.hi[2];
  bool inside;
  float w = lw.valid ? unnormalised_weight(x, lw.x, lw.y, lw.z, inside) : 0.0f;
  const float ws = wave_sum(w);
  if (ws > 0.0f) w = w / ws;
  lw.w = w;
  lw.wsum = wave_sum(w);
  return lw;
}

// Particle state layout (double[P][FUNLBM_PSTRIDE]).
__device__ inline double* pst(double* ps, int i) { return ps + (int64_t)i * FUNLBM_PSTRIDE; }
__device__ inline const double* pst(const double* ps, int i) { return ps + (int64_t)i * FUNLBM_PSTRIDE; }

// ---------------------------------------------------------------------------
// Interpolation + direct forcing (src/funlbm/lbm/lbm3d.py:66-100):
//   s_u = sum_nodes sum_d u_d w (the three components summed into one scalar),
//   s_rho = sum w rho, lu = cu + cw x (lx - cx), lF = 2 s_rho (lu - s_u)/dt.
// FROM_STATE: u', rho' are taken by pulling the post-stream populations out of
// the fused state and forming the moments with FOL_old; otherwise from u/rho.
// ---------------------------------------------------------------------------
template <class L, class S, bool FROM_STATE>
__global__ __launch_bounds__(kIbmBlock) void k_interp(
    KParams p, const typename S::elem* __restrict__ state, const float* __restrict__ fol_old,
    const float* __restrict__ rho_face, const float* __restrict__ ufield,
    const float* __restrict__ rhofield, int nm, const float* __restrict__ lx,
    const int* __restrict__ pid, const double* __restrict__ ps, float dt, int check_sum,
    float* __restrict__ lu, float* __restrict__ lF, int* __restrict__ ctl) {
  using T = typename S::T;
  const int m = blockIdx.x * (kIbmBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (m >= nm) return;
  const float x[3] = {lx[3 * m], lx[3 * m + 1], lx[3 * m + 2]};
  const LaneW lw = lane_weight(p, x, lane);
  float su = 0.0f, sr = 0.0f;
  if (lw.valid) {
    const int64_t node = lin(p, lw.x, lw.y, lw.z);
    float uu[3], rr;
    if constexpr (FROM_STATE) {
      const Node c{lw.x, lw.y, lw.z};
      T f[L::Q];
      pull_node<L, S>(p, state, c, node, rho_face, f);
      T F[3];
#pragma unroll
      for (int d = 0; d < 3; ++d)
        F[d] = (T)(fol_old ? fol_old[(int64_t)d * p.n + node] + p.g[d] : p.g[d]);
      T r, u[3];
      moments<L, T>(p, c, f, F, r, u);
      rr = (float)r;
#pragma unroll
      for (int d = 0; d < 3; ++d) uu[d] = (float)u[d];
    } else {
      rr = rhofield[node];
#pragma unroll
      for (int d = 0; d < 3; ++d) uu[d] = ufield[(int64_t)d * p.n + node];
    }
    su = (uu[0] * lw.w + uu[1] * lw.w) + uu[2] * lw.w;
    sr = rr * lw.w;
  }
  su = wave_sum(su);
  sr = wave_sum(sr);
  if (lane == 0) {
    if (check_sum && fabsf(lw.wsum - 1.0f) > 1e-3f) atomicOr(&ctl[1], 1);
    const double* P = pst(ps, pid[m]);
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
}

// ---------------------------------------------------------------------------
// Wall spring (src/funlbm/lbm/lbm3d.py:121-167), one block per particle.
// Extreme markers are the first arg-min / arg-max per axis (torch's tie rule);
// d = max(k0 (1 - |coord - bound|/(2 dx)), 0), bounds [0,0,0,size]; the six
// rows lF[ext_j] + n_j d_j are gathered before any is written, then written in
// order j = 0..5 so a duplicate index keeps the last row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kIbmBlock) void k_wall(KParams p, const int* __restrict__ offsets,
                                                    const float* __restrict__ lx,
                                                    float* __restrict__ lF, float dx) {
  const int part = blockIdx.x;
  const int b = offsets[part], e = offsets[part + 1];
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
        const bool better = (j < 3) ? (v < cv || (v == cv && i < ci)) : (v > cv || (v == cv && i < ci));
        if (better) {
          sv[j][threadIdx.x] = v;
          si[j][threadIdx.x] = i;
        }
      }
    }
    __syncthreads();
  }
  if (threadIdx.x != 0 || e <= b) return;
  const float bounds[6] = {0.0f, 0.0f, 0.0f, (float)(p.nx - 1), (float)(p.ny - 1), (float)(p.nz - 1)};
  const float normals[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
  float dist[6];
  float dmax = 0.0f;
  int ext[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    ext[j] = si[j][0];
    const float c = lx[3 * ext[j] + (j % 3)];
    float d = 300.0f * (1.0f - fabsf(c - bounds[j]) / (2.0f * dx));
    d = fmaxf(d, 0.0f);
    dist[j] = d;
    dmax = fmaxf(dmax, d);
  }
  if (dmax == 0.0f) return;
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

// ---------------------------------------------------------------------------
// Rigid-body update (src/funlbm/particle/base/base.py:120-151) in double:
// cF = sum(-lF lm) + (1 - rho_p/rho_f) V g; cu += cF/V dt; cx += cu dt;
// cT = -sum((lx - cx) x lF) lm; cw += 0.1 cT dt / I.  One block per particle,
// fixed-order tree reductions.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kIbmBlock) void k_forces(const int* __restrict__ offsets,
                                                      const float* __restrict__ lx,
                                                      const float* __restrict__ lF,
                                                      const float* __restrict__ lm,
                                                      double* __restrict__ ps, double dt,
                                                      double rho_f) {
  const int part = blockIdx.x;
  const int b = offsets[part], e = offsets[part + 1];
  double* P = pst(ps, part);
  __shared__ double red[6][kIbmBlock];
  __shared__ double cxn[3];
  // pass 1: force
  double fs[3] = {0, 0, 0};
  for (int m = b + threadIdx.x; m < e; m += kIbmBlock)
#pragma unroll
    for (int a = 0; a < 3; ++a) fs[a] += -(double)lF[3 * m + a] * (double)lm[m];
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
  // pass 2: torque about the updated centre
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
// lx = float32(R lag) + float32(cx).  One block per particle.
__global__ __launch_bounds__(kIbmBlock) void k_pose(const int* __restrict__ offsets,
                                                    const float* __restrict__ lag,
                                                    float* __restrict__ lx, double* __restrict__ ps,
                                                    int advance) {
  const int part = blockIdx.x;
  const int b = offsets[part], e = offsets[part + 1];
  double* P = pst(ps, part);
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
}

// ---------------------------------------------------------------------------
// Spreading (src/funlbm/lbm/lbm3d.py:102-119): FOL := 0, then for every marker
// in order FOL[stencil] = w lF lm, overwriting.  A node therefore ends up with
// the value of the highest-numbered marker whose stencil box covers it.
//   prep   zero the stencils this buffer received last time (its old support)
//          and raise owner[node] to (stamp, marker) with an integer atomicMax
//   write  the owning marker of each node writes w lF lm
//   bump   advance the stamp, so owner[] never needs clearing
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kIbmBlock) void k_spread_prep(KParams p, int nm,
                                                           const float* __restrict__ lx,
                                                           const int* __restrict__ clear_origin,
                                                           float* __restrict__ fol,
                                                           unsigned long long* __restrict__ owner,
                                                           const int* __restrict__ ctl) {
  const int m = blockIdx.x * (kIbmBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (m >= nm) return;
  const int i = lane >> 4, j = (lane >> 2) & 3, k = lane & 3;
  if (clear_origin) {
    const int ox = clear_origin[3 * m] + i, oy = clear_origin[3 * m + 1] + j,
              oz = clear_origin[3 * m + 2] + k;
    if (ox < p.nx && oy < p.ny && oz < p.nz) {
      const int64_t node = lin(p, ox, oy, oz);
#pragma unroll
      for (int d = 0; d < 3; ++d) fol[(int64_t)d * p.n + node] = 0.0f;
    }
  }
  const float x[3] = {lx[3 * m], lx[3 * m + 1], lx[3 * m + 2]};
  const Stencil s = stencil_of(p, x);
  const int nx_ = s.lo[0] + i, ny_ = s.lo[1] + j, nz_ = s.lo[2] + k;
  if (nx_ < s.hi[0] && ny_ < s.hi[1] && nz_ < s.hi[2]) {
    const unsigned long long key =
        ((unsigned long long)(unsigned)(ctl[0] + 1) << 32) | (unsigned long long)(unsigned)m;
    atomicMax(&owner[lin(p, nx_, ny_, nz_)], key);
  }
}

__global__ __launch_bounds__(kIbmBlock) void k_spread_write(
    KParams p, int nm, const float* __restrict__ lx, const float* __restrict__ lF,
    const float* __restrict__ lm, float* __restrict__ fol,
    const unsigned long long* __restrict__ owner, int* __restrict__ origin, const int* __restrict__ ctl) {
  const int m = blockIdx.x * (kIbmBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (m >= nm) return;
  const float x[3] = {lx[3 * m], lx[3 * m + 1], lx[3 * m + 2]};
  const LaneW lw = lane_weight(p, x, lane);
  if (lane == 0 && origin) {
    const Stencil s = stencil_of(p, x);
    origin[3 * m] = s.lo[0];
    origin[3 * m + 1] = s.lo[1];
    origin[3 * m + 2] = s.lo[2];
  }
  if (!lw.valid) return;
  const int64_t node = lin(p, lw.x, lw.y, lw.z);
  const unsigned long long key =
      ((unsigned long long)(unsigned)(ctl[0] + 1) << 32) | (unsigned long long)(unsigned)m;
  if (owner[node] != key) return;
  const float mm = lm[m];
#pragma unroll
  for (int d = 0; d < 3; ++d) fol[(int64_t)d * p.n + node] = (lw.w * lF[3 * m + d]) * mm;
}

__global__ void k_bump(int* ctl) { ctl[0] += 1; }

static int wave_grid(int nm) { return (nm + (kIbmBlock / kWave) - 1) / (kIbmBlock / kWave); }

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
                      uint64_t* owner, int32_t* ctl, void* stream) {
  KParams p;
  int rc = make_params(lat, &p);
  if (rc) return rc;
  if (n_markers < 0) return set_error(FUNLBM_EINVAL, "ibm_spread: negative marker count");
  if (!fol || !owner || !ctl) return set_error(FUNLBM_EINVAL, "ibm_spread: NULL array");
  if (n_markers > 0 && (!lx || !lF || !lm)) return set_error(FUNLBM_EINVAL, "ibm_spread: NULL array");
  hipStream_t s = (hipStream_t)stream;
  if (n_markers > 0) {
    hipLaunchKernelGGL(k_spread_prep, dim3(wave_grid(n_markers)), dim3(kIbmBlock), 0, s, p, n_markers,
                       lx, clear_origin, fol, (unsigned long long*)owner, ctl);
    if ((rc = check_launch("ibm_spread"))) return rc;
    hipLaunchKernelGGL(k_spread_write, dim3(wave_grid(n_markers)), dim3(kIbmBlock), 0, s, p, n_markers,
                       lx, lF, lm, fol, (const unsigned long long*)owner, origin, ctl);
    if ((rc = check_launch("ibm_spread"))) return rc;
  }
  hipLaunchKernelGGL(k_bump, dim3(1), dim3(1), 0, s, ctl);
  return check_launch("ibm_spread");
}

}  // extern "C"
