// Records gfx950's stochastic conversions (v_cvt_sr_f16_f32,
// v_cvt_sr_bf16_f32) for crafted and pseudo-random (x, r) pairs: writes
// int32 n, float x[n], uint32 r[n], then the f16 and the bf16 results (one
// uint32 each, the value in the low half) to argv[1] (default srprobe.bin).
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdint.h>
#include <string.h>
#include <vector>
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

__global__ void k(const float* x, const unsigned* r, uint32_t* o16, uint32_t* ob, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  f16x2 a = {0, 0};
  a = __builtin_amdgcn_cvt_sr_f16_f32(a, x[i], r[i], false);
  bf16x2 b = {0, 0};
  b = __builtin_amdgcn_cvt_sr_bf16_f32(b, x[i], r[i], false);
  uint32_t ua, ub;
  memcpy(&ua, &a, 4);
  memcpy(&ub, &b, 4);
  o16[i] = ua;
  ob[i] = ub;
}

static uint32_t lcg(uint64_t& s) { s = s * 6364136223846793005ull + 1442695040888963407ull; return (uint32_t)(s >> 32); }

int main(int argc, char** argv) {
  std::vector<float> xs;
  std::vector<uint32_t> rs;
  const float base[] = {1.0f, 1.00048828125f, 0.3f, -0.3f, 1e-3f, -1e-3f, 6.1e-5f, 3e-6f, -3e-6f, 6e-8f, 1e-9f,
                        65504.f, 65519.f, 65520.f, 70000.f, -70000.f, 0.0f, -0.0f, 1.5f, 123.456f, 2.5e38f};
  const uint32_t rr[] = {0u, 1u, 0xFFFu, 0x1000u, 0x1FFFu, 0x2000u, 0x7FFFu, 0x8000u, 0xFFFFu, 0x10000u,
                         0x7FFFFu, 0x80000u, 0xFFF80000u, 0xFFFFE000u, 0xFFFF0000u, 0xFFFFFFFFu, 0x80000000u};
  for (float b : base)
    for (uint32_t r : rr) { xs.push_back(b); rs.push_back(r); }
  uint64_t s = 12345;
  for (int i = 0; i < 200000; ++i) {
    uint32_t bits = lcg(s);
    // exponents spread over f16 normal, subnormal and tiny ranges
    uint32_t e = 90 + (lcg(s) % 60);
    bits = (bits & 0x807FFFFFu) | (e << 23);
    float f;
    memcpy(&f, &bits, 4);
    xs.push_back(f);
    rs.push_back(lcg(s));
  }
  const int n = (int)xs.size();
  float* dx; uint32_t *dr, *d16, *db;
  if (hipMalloc(&dx, n * 4) || hipMalloc(&dr, n * 4) || hipMalloc(&d16, n * 4) || hipMalloc(&db, n * 4)) return 2;
  hipMemcpy(dx, xs.data(), n * 4, hipMemcpyHostToDevice);
  hipMemcpy(dr, rs.data(), n * 4, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k, dim3((n + 255) / 256), dim3(256), 0, 0, dx, dr, d16, db, n);
  if (hipDeviceSynchronize() != hipSuccess) return 3;
  std::vector<uint32_t> o16(n), ob(n);
  hipMemcpy(o16.data(), d16, n * 4, hipMemcpyDeviceToHost);
  hipMemcpy(ob.data(), db, n * 4, hipMemcpyDeviceToHost);
  FILE* fp = fopen(argc > 1 ? argv[1] : "srprobe.bin", "wb");
  if (!fp) return 4;
  fwrite(&n, 4, 1, fp);
  fwrite(xs.data(), 4, n, fp);
  fwrite(rs.data(), 4, n, fp);
  fwrite(o16.data(), 4, n, fp);
  fwrite(ob.data(), 4, n, fp);
  fclose(fp);
  printf("srprobe wrote %d cases\n", n);
  return 0;
}
