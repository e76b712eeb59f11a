// Surface markers of ellipsoids on the GPU: replaces the init-time
// generate_uniform_points_on_ellipsoid / find_intersection
// (src/funlbm/particle/base/ellipsoid.py:10-95) that Sphere._init and
// Ellipsoid._init call (src/funlbm/particle/base/sphere.py:15-53,
// ellipsoid.py:107-143).
//
// Unit cubes of edge dx tile the box [-s-2dx, s+2dx) per semi-axis s (the
// axis values of numpy.arange: v0 = start, v1 = start + dx,
// vi = start + i ((start + dx) - start)).  A cube whose eight vertices are
// neither all inside (x^2/a^2 + y^2/b^2 + z^2/c^2 - 1 < 0) nor all outside
// contributes the intersection point of every vertex pair of differing side
// (the 28 pairs in order), and the centroid of those points is a marker when
// there are at least 3.  Markers come out in cube order (x outermost).
//
// The reference finds each intersection with scipy's fsolve started at
// t = 0.5; on a segment whose end points lie on opposite sides of the convex
// quadric the quadratic in t has one root in [0, 1], computed here in closed
// form with the cancellation-free pair q = -(B + sign(B) sqrt(B^2-4AC))/2,
// t in {q/A, C/q}, taking the root nearer to [0, 1].  All of it in double,
// with the operation order of the numpy restatement, so the float markers are
// bit-identical to it.
//
// One workgroup per shape; the cubes are visited in chunks of the block size
// and compacted in order with a wave ballot and a per-chunk block prefix.

#include "capi_util.h"

namespace funlbm {

namespace {

constexpr int kMarkerBlock = 256;
constexpr int kMarkerWaves = kMarkerBlock / 64;
// Axis points beyond which a shape is rejected (a semi-axis of about 510 dx;
// the cube count then stays below 2^31).
constexpr double kMaxAxisPoints = 1024.0;

struct Axis {
  double start, step1, delta;
  int n;
};

__device__ inline Axis make_axis(double s, double dx) {
  Axis ax;
  ax.start = -s - 2.0 * dx;
  const double stop = s + 2.0 * dx;
  const double len = ceil((stop - ax.start) / dx);
  ax.n = (len >= 2.0 && len <= kMaxAxisPoints) ? (int)len : -1;
  ax.step1 = ax.start + dx;
  ax.delta = ax.step1 - ax.start;
  return ax;
}

__device__ inline double axis_value(const Axis& ax, int i) {
  if (i == 0) return ax.start;
  if (i == 1) return ax.step1;
  return ax.start + (double)i * ax.delta;
}

// In [0, 1] root t of q(p1 + t (p2 - p1)) = 0, q = sum x^2 inv2 - 1.
__device__ inline double segment_root(const double p1[3], const double d[3], const double inv2[3]) {
  const double A = (d[0] * d[0] * inv2[0] + d[1] * d[1] * inv2[1]) + d[2] * d[2] * inv2[2];
  const double B = 2.0 * ((p1[0] * d[0] * inv2[0] + p1[1] * d[1] * inv2[1]) + p1[2] * d[2] * inv2[2]);
  const double C = ((p1[0] * p1[0] * inv2[0] + p1[1] * p1[1] * inv2[1]) + p1[2] * p1[2] * inv2[2]) - 1.0;
  const double disc = sqrt(fmax(B * B - 4.0 * A * C, 0.0));
  const double qv = -0.5 * (B + copysign(disc, B));
  const double r1 = qv / A;
  const double r2 = C / qv;
  double d1 = fmax(fmax(-r1, r1 - 1.0), 0.0);
  double d2 = fmax(fmax(-r2, r2 - 1.0), 0.0);
  if (!isfinite(r1)) d1 = INFINITY;
  if (!isfinite(r2)) d2 = INFINITY;
  return d1 <= d2 ? r1 : r2;
}

__global__ __launch_bounds__(kMarkerBlock) void k_surface_markers(const double* __restrict__ semi, double dx,
                                                                  int32_t cap, float* __restrict__ out,
                                                                  int32_t* __restrict__ counts) {
  __shared__ int wave_total[kMarkerWaves];
  const int shape = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const double a = semi[3 * shape + 0], b = semi[3 * shape + 1], c = semi[3 * shape + 2];
  const double s[3] = {a, b, c};
  bool ok = dx > 0.0 && isfinite(dx);
  Axis ax[3];
  double inv2[3], sq[3];
  for (int d = 0; d < 3; ++d) {
    ok = ok && s[d] > 0.0 && isfinite(s[d]);
    ax[d] = make_axis(s[d], dx);
    ok = ok && ax[d].n > 0;
    sq[d] = s[d] * s[d];
    inv2[d] = 1.0 / sq[d];
  }
  if (!ok) {
    if (tid == 0) counts[shape] = -1;
    return;
  }
  const int cy = ax[1].n - 1, cz = ax[2].n - 1;
  const int n_cubes = (ax[0].n - 1) * cy * cz;
  float* dst = out + (int64_t)shape * cap * 3;
  int base = 0;
  for (int chunk = 0; chunk < n_cubes; chunk += kMarkerBlock) {
    const int cube = chunk + tid;
    bool keep = false;
    double m[3] = {0.0, 0.0, 0.0};
    if (cube < n_cubes) {
      const int k = cube % cz;
      const int j = (cube / cz) % cy;
      const int i = cube / (cz * cy);
      const double x0 = axis_value(ax[0], i), x1 = axis_value(ax[0], i + 1);
      const double y0 = axis_value(ax[1], j), y1 = axis_value(ax[1], j + 1);
      const double z0 = axis_value(ax[2], k), z1 = axis_value(ax[2], k + 1);
      // vertex v = (x bit 2, y bit 1, z bit 0); coordinates by select, no arrays
#define VX(v) (((v)&4) ? x1 : x0)
#define VY(v) (((v)&2) ? y1 : y0)
#define VZ(v) (((v)&1) ? z1 : z0)
      unsigned inside = 0;
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const double val = ((VX(v) * VX(v) / sq[0] + VY(v) * VY(v) / sq[1]) + VZ(v) * VZ(v) / sq[2]) - 1.0;
        inside |= (val < 0.0 ? 1u : 0u) << v;
      }
      if (inside != 0u && inside != 0xFFu) {
        int count = 0;
        for (int v1 = 0; v1 < 8; ++v1) {
          for (int v2 = v1 + 1; v2 < 8; ++v2) {
            if (((inside >> v1) ^ (inside >> v2)) & 1u) {
              const double p1[3] = {VX(v1), VY(v1), VZ(v1)};
              const double d[3] = {VX(v2) - p1[0], VY(v2) - p1[1], VZ(v2) - p1[2]};
              const double t = segment_root(p1, d, inv2);
              m[0] = m[0] + (p1[0] + t * d[0]);
              m[1] = m[1] + (p1[1] + t * d[1]);
              m[2] = m[2] + (p1[2] + t * d[2]);
              ++count;
            }
          }
        }
#undef VX
#undef VY
#undef VZ
        keep = count >= 3;
        if (keep) {
          m[0] = m[0] / (double)count;
          m[1] = m[1] / (double)count;
          m[2] = m[2] / (double)count;
        }
      }
    }
    // in-order compaction: prefix within the wave, then over the waves
    const unsigned long long ballot = __ballot(keep);
    const int before = __popcll(ballot & ((1ull << lane) - 1ull));
    if (lane == 0) wave_total[wave] = __popcll(ballot);
    __syncthreads();
    int offset = base + before, total = 0;
    for (int w = 0; w < kMarkerWaves; ++w) {
      offset += w < wave ? wave_total[w] : 0;
      total += wave_total[w];
    }
    if (keep && offset < cap) {
      dst[3 * offset + 0] = (float)m[0];
      dst[3 * offset + 1] = (float)m[1];
      dst[3 * offset + 2] = (float)m[2];
    }
    base += total;
    __syncthreads();
  }
  if (tid == 0) counts[shape] = base;
}

}  // namespace

}  // namespace funlbm

using namespace funlbm;

extern "C" {

int funlbm_surface_markers(int32_t n_shapes, const double* semi, double dx, int32_t cap, float* markers,
                           int32_t* counts, void* stream) {
  if (n_shapes <= 0) return n_shapes == 0 ? 0 : set_error(FUNLBM_EINVAL, "surface_markers: negative count");
  if (cap < 0) return set_error(FUNLBM_EINVAL, "surface_markers: negative capacity");
  if (!(dx > 0.0) || !isfinite(dx)) return set_error(FUNLBM_EINVAL, "surface_markers: dx must be positive");
  if (!semi || !counts || (cap > 0 && !markers))
    return set_error(FUNLBM_EINVAL, "surface_markers: NULL array");
  if ((int64_t)n_shapes * cap * 3 > ((int64_t)1 << 40))
    return set_error(FUNLBM_ETOOBIG, "surface_markers: output too large");
  hipLaunchKernelGGL(k_surface_markers, dim3(n_shapes), dim3(kMarkerBlock), 0, (hipStream_t)stream, semi, dx,
                     cap, markers, counts);
  return check_launch("surface_markers");
}

}  // extern "C"
