// Host-side helpers shared by the C-ABI entry points: thread-local error
// text, argument validation and launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "../../include/funlbm.h"
#include "lbm_common.h"

namespace funlbm {

inline char* error_buffer() {
  static thread_local char buf[512] = {0};
  return buf;
}

inline const char* last_error_message() { return error_buffer(); }

inline int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(error_buffer(), 512, fmt, ap);
  va_end(ap);
  return code;
}

inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error((int)e, "%s: %s", what, hipGetErrorString(e));
  return 0;
}

inline int storage_elem_bytes(int storage) {
  switch (storage) {
    case FUNLBM_STORE_F32: return 4;
    case FUNLBM_STORE_F16: return 2;
    case FUNLBM_STORE_BF16: return 2;
    case FUNLBM_STORE_F64: return 8;
    default: return 0;
  }
}

// Multiplier and shift with floor(n / d) = (n * m) >> s for all n < 2^31
// (round-up method of Granlund and Montgomery: s = 31 + ceil(log2 d),
// m = floor(2^s / d) + 1 < 2^32).
inline bool magic_div(int d, uint32_t* m, int* s) {
  if (d < 1) return false;
  int l = 0;
  while ((1LL << l) < d) ++l;
  const int sh = 31 + l;
  const unsigned __int128 two = (unsigned __int128)1 << sh;
  const unsigned __int128 mm = two / (unsigned __int128)d + 1;
  if (mm >= ((unsigned __int128)1 << 32)) return false;
  *m = (uint32_t)mm;
  *s = sh;
  return true;
}

// Tile shift of the fused-state layout (state_index in lbm_common.h): tiles
// of 4096 nodes from 2^14 nodes up, of 256 below (one kernel block; keeps the
// two slack tiles of a tiny lattice small).  On the MI355X a 19-plane pull
// and store over 4096-node tiles measured 5.97 TB/s against 5.61 TB/s for
// planes of n elements padded to avoid power-of-two strides.  Large tiles also
// keep a wavefront's shifted sources inside one tile more often on lattices
// whose row lengths are not multiples of 64: 256 settling cases of 32^3 to
// 96^3 nodes stepped 6 % faster with 4096-node tiles than with 256-node ones.
inline int state_tile_shift(int64_t n) { return n >= (1 << 14) ? 12 : 8; }

// Elements of one fused-state buffer: the tiles of n nodes plus two slack
// tiles, Q planes each.
inline int64_t state_elems(int q, int64_t n, int ts) {
  const int64_t T = (int64_t)1 << ts;
  return ((n + T - 1) / T + 2) * T * q;
}

// The pull-source table of KParams (soff, mixmask) for lattice L; false if
// an offset does not fit 32 bits.
template <class L>
inline bool fill_pull_table(KParams* p, int esz) {
  const int64_t T = (int64_t)1 << p->ts;
  for (int k = 0; k < 27; ++k) p->soff[k] = 0;
  p->mixmask = 0;
  for (int k = 0; k < L::Q; ++k) {
    const int64_t sk = (int64_t)L::E[k][0] * p->sxy + (int64_t)L::E[k][1] * p->nz + L::E[k][2];
    const int64_t qk = sk >= 0 ? sk / T : -((-sk + T - 1) / T);  // floor
    const int64_t rk = sk - qk * T;
    const int64_t so = ((int64_t)L::Q - qk * L::Q + k) * T * esz - rk * esz;
    if (so > 0x7fffffffLL || so < -0x7fffffffLL) return false;
    p->soff[k] = (int32_t)so;
    if (rk % 64 != 0) p->mixmask |= (uint64_t)1 << (rk / 64);  // waves with l0 < rk < l0 + 64
  }
  return true;
}

// Validates a lattice descriptor and fills the kernel parameters.
inline int make_params(const funlbm_lattice* lat, KParams* p) {
  if (!lat) return set_error(FUNLBM_EINVAL, "lattice descriptor is NULL");
  if (lat->q != 13 && lat->q != 15 && lat->q != 19 && lat->q != 27)
    return set_error(FUNLBM_EUNSUPPORTED_LATTICE, "unknown lattice D3Q%d", lat->q);
  if (lat->nx < 1 || lat->ny < 1 || lat->nz < 1)
    return set_error(FUNLBM_EINVAL, "bad grid %d x %d x %d", lat->nx, lat->ny, lat->nz);
  if (storage_elem_bytes(lat->storage) == 0)
    return set_error(FUNLBM_EINVAL, "unknown storage code %d", lat->storage);
  if (!(lat->tau > 0.0f) || !isfinite(lat->tau))
    return set_error(FUNLBM_EINVAL, "tau must be positive and finite, got %g", (double)lat->tau);
  const bool dev = lat->storage == FUNLBM_STORE_F16 || lat->storage == FUNLBM_STORE_BF16;
  if (dev && !(lat->dev_scale > 0.0f))
    return set_error(FUNLBM_EINVAL, "deviation storage needs dev_scale > 0");
  p->nx = lat->nx;
  p->ny = lat->ny;
  p->nz = lat->nz;
  p->n = (int64_t)lat->nx * lat->ny * lat->nz;
  if (p->n > 0x7fffffffLL)
    return set_error(FUNLBM_ETOOBIG, "%lld nodes exceed the 2^31 - 1 node limit", (long long)p->n);
  p->sxy = lat->ny * lat->nz;
  p->has_moving = 0;
  p->has_neq = 0;
  for (int f = 0; f < 6; ++f) {
    const int bc = lat->bc[f];
    const bool neq_ok = bc == FUNLBM_BC_NON_EQUILIBRIUM && f < 2;  // the reference's x faces only
    if (bc != FUNLBM_BC_PERIODICAL && bc != FUNLBM_BC_WALL && bc != FUNLBM_BC_WALL_WITH_SPEED && !neq_ok)
      return set_error(FUNLBM_EUNSUPPORTED_BC, "boundary code %d on face %d is not implemented",
                       bc, f);
    p->bc[f] = bc;
    p->has_moving |= (bc == FUNLBM_BC_WALL_WITH_SPEED);
    p->has_neq |= neq_ok;
    for (int d = 0; d < 3; ++d) p->uw[f][d] = lat->uw[f][d];
  }
  if (p->has_neq && lat->nx < 3)
    return set_error(FUNLBM_EINVAL, "NON_EQUILIBRIUM faces need nx >= 3, got %d", lat->nx);
  p->uzero = lat->uzero_mask & 0x3f;
  p->tau = lat->tau;
  // the reference's float32 arithmetic on the host: correctly rounded
  p->rtau = 1.0f / lat->tau;
  p->omega = 1.0f - 1.0f / (2.0f * lat->tau);
  p->tau_d = (double)lat->tau;
  p->rtau_d = 1.0 / p->tau_d;
  p->omega_d = 1.0 - 1.0 / (2.0 * p->tau_d);
  if (!magic_div(p->nz, &p->mz, &p->sz) || !magic_div(p->ny, &p->my, &p->sy))
    return set_error(FUNLBM_EINVAL, "grid extent without a 32-bit division multiplier");
  if (dev && lat->rho0 != 1.0f)
    return set_error(FUNLBM_EINVAL, "rho0 must be 1 (the deviation base of the 16-bit states is w_k), got %g",
                     (double)lat->rho0);
  if (lat->rounding != FUNLBM_ROUND_NEAREST && lat->rounding != FUNLBM_ROUND_STOCHASTIC)
    return set_error(FUNLBM_EINVAL, "unknown rounding code %d", lat->rounding);
  if (lat->rounding == FUNLBM_ROUND_STOCHASTIC && !dev)
    return set_error(FUNLBM_EINVAL, "stochastic rounding applies to F16/BF16 storage only");
  p->sr = lat->rounding == FUNLBM_ROUND_STOCHASTIC;
  if (lat->delta != FUNLBM_DELTA_COSINE && lat->delta != FUNLBM_DELTA_PESKIN4)
    return set_error(FUNLBM_EINVAL, "unknown immersed-boundary kernel code %d", lat->delta);
  p->delta = lat->delta;
  p->seed = lat->seed;
  p->seed_mix = mix32(lat->seed);
  p->rho0 = lat->rho0;
  p->scale = dev ? lat->dev_scale : 1.0f;
  p->inv_scale = dev ? 1.0f / lat->dev_scale : 1.0f;
  p->has_body = 0;
  for (int d = 0; d < 3; ++d) {
    p->g[d] = lat->body_force[d];
    p->has_body |= (lat->body_force[d] != 0.0f);
  }
  p->ts = state_tile_shift(p->n);
  {
    const int esz = storage_elem_bytes(lat->storage);
    bool ok;
    switch (lat->q) {
      case 13: ok = fill_pull_table<D3Q13>(p, esz); break;
      case 15: ok = fill_pull_table<D3Q15>(p, esz); break;
      case 19: ok = fill_pull_table<D3Q19>(p, esz); break;
      default: ok = fill_pull_table<D3Q27>(p, esz); break;
    }
    if (!ok)
      return set_error(FUNLBM_ETOOBIG, "grid cross-section %d x %d too large for 32-bit pull offsets",
                       lat->ny, lat->nz);
  }
  return 0;
}

// The fused state paths (fused step, materialise, statistics, batched cases,
// interpolation from the state) do not implement NON_EQUILIBRIUM faces.
inline int reject_neq(const KParams& p, const char* what) {
  if (p.has_neq)
    return set_error(FUNLBM_EUNSUPPORTED_BC, "%s: NON_EQUILIBRIUM faces run on the explicit path", what);
  return 0;
}

template <class Fn>
inline int dispatch_lattice(int q, Fn&& fn) {
  switch (q) {
    case 13: return fn(D3Q13{});
    case 15: return fn(D3Q15{});
    case 19: return fn(D3Q19{});
    case 27: return fn(D3Q27{});
    default: return set_error(FUNLBM_EUNSUPPORTED_LATTICE, "unknown lattice D3Q%d", q);
  }
}

template <class Fn>
inline int dispatch_storage(int storage, Fn&& fn) {
  switch (storage) {
    case FUNLBM_STORE_F32: return fn(StoreF32{});
    case FUNLBM_STORE_F16: return fn(StoreF16{});
    case FUNLBM_STORE_BF16: return fn(StoreBF16{});
    case FUNLBM_STORE_F64: return fn(StoreF64{});
    default: return set_error(FUNLBM_EINVAL, "unknown storage code %d", storage);
  }
}

}  // namespace funlbm
