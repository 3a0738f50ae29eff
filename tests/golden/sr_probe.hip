// Generator of tests/golden/sr_vectors.npz: runs gfx950's stochastic-rounding
// conversions (v_cvt_sr_bf16_f32, v_cvt_sr_f16_f32) on 2^20 (x, r) pairs and
// writes x, r and both outputs to the file named on the command line (default
// sr_probe.bin); the fixture keeps
// 2048 pairs of each of the 8 input classes (NaN inputs dropped).
// Build: hipcc --offload-arch=gfx950 -O2 -o sr_probe sr_probe.hip
// One thread per pair, grid = ceil(N/256) blocks, guarded by i < N.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <vector>
#include <cstring>
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
__global__ void k(const float* x, const unsigned* r, unsigned short* ob, unsigned short* oh, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bf16x2 b = {};
  b = __builtin_amdgcn_cvt_sr_bf16_f32(b, x[i], r[i], false);
  f16x2 h = {};
  h = __builtin_amdgcn_cvt_sr_f16_f32(h, x[i], r[i], false);
  ob[i] = __builtin_bit_cast(unsigned short, b[0]);
  oh[i] = __builtin_bit_cast(unsigned short, h[0]);
}
static uint32_t s = 12345;
static uint32_t xr() { s ^= s << 13; s ^= s >> 17; s ^= s << 5; return s; }
int main(int argc, char** argv) {
  const int n = 1 << 20;
  std::vector<float> x(n); std::vector<unsigned> r(n);
  for (int i = 0; i < n; ++i) {
    uint32_t u = xr();
    int cls = i % 8;
    float v;
    if (cls < 3) { float m = (xr() >> 8) * (1.0f / 16777216.0f); v = (m - 0.5f) * 0.1f; }          // |x| < 0.05
    else if (cls == 3) { float m = (xr() >> 8) * (1.0f / 16777216.0f); v = (m - 0.5f) * 2e-4f; }   // f16 subnormal range
    else if (cls == 4) { uint32_t b = (u & 0x807fffffu) | ((uint32_t)(100 + (xr() % 60)) << 23); std::memcpy(&v, &b, 4); } // wide exponents
    else if (cls == 5) { uint32_t b = (u & 0x807fffffu) | (0x8eu << 23); std::memcpy(&v, &b, 4); }  // near 65504
    else { std::memcpy(&v, &u, 4); if (v != v) v = 1.0f; }                                         // any bits
    x[i] = v; r[i] = xr();
  }
  float* dx; unsigned* dr; unsigned short *db, *dh;
  if (hipMalloc(&dx, n * 4) || hipMalloc(&dr, n * 4) || hipMalloc(&db, n * 2) || hipMalloc(&dh, n * 2)) { printf("alloc fail\n"); return 1; }
  hipMemcpy(dx, x.data(), n * 4, hipMemcpyHostToDevice); hipMemcpy(dr, r.data(), n * 4, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k, dim3((n + 255) / 256), dim3(256), 0, 0, dx, dr, db, dh, n);
  std::vector<unsigned short> ob(n), oh(n);
  if (hipMemcpy(ob.data(), db, n * 2, hipMemcpyDeviceToHost) != hipSuccess) { printf("fail\n"); return 1; }
  hipMemcpy(oh.data(), dh, n * 2, hipMemcpyDeviceToHost);
  FILE* f = fopen(argc > 1 ? argv[1] : "sr_probe.bin", "wb");
  if (!f) { printf("cannot open output\n"); return 1; }
  fwrite(x.data(), 4, n, f); fwrite(r.data(), 4, n, f); fwrite(ob.data(), 2, n, f); fwrite(oh.data(), 2, n, f); fclose(f);
  printf("sr probe ok n=%d\n", n);
  return 0;
}
