// valu_peak.hip -- issue-rate probe for the VALU instructions the engine uses.
// Long independent chains of one instruction kind per lane, written
// as inline asm so the compiler cannot fold them. ALU only, no memory in asm.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#define CH 8
#define BODY(STMT) for (int i = 0; i < iters; ++i) { _Pragma("unroll") for (int u = 0; u < 16; ++u) { _Pragma("unroll") for (int c = 0; c < CH; ++c) { STMT; } } }
__global__ __launch_bounds__(256) void k_mad(const uint32_t* in, uint64_t* out, int iters) {
  uint32_t x = in[threadIdx.x & 63] | 1u; uint64_t acc[CH];
  for (int c = 0; c < CH; ++c) acc[c] = x + c;
  BODY(asm volatile("v_mad_u64_u32 %0, s[40:41], %1, %1, %0" : "+v"(acc[c]) : "v"(x) : "s40", "s41"));
  uint64_t s = 0; for (int c = 0; c < CH; ++c) s ^= acc[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_fma(const uint32_t* in, uint64_t* out, int iters) {
  double x = (double)(in[threadIdx.x & 63] | 1u) * 1e-9; double acc[CH];
  for (int c = 0; c < CH; ++c) acc[c] = c;
  BODY(asm volatile("v_fma_f64 %0, %1, %1, %0" : "+v"(acc[c]) : "v"(x)));
  double s = 0; for (int c = 0; c < CH; ++c) s += acc[c];
  out[blockIdx.x * 256 + threadIdx.x] = (uint64_t)s;
}
__global__ __launch_bounds__(256) void k_add64(const uint32_t* in, uint64_t* out, int iters) {
  uint64_t x = in[threadIdx.x & 63] | 1u; uint64_t acc[CH];
  for (int c = 0; c < CH; ++c) acc[c] = x + c;
  BODY(asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(acc[c]) : "v"(x)));
  uint64_t s = 0; for (int c = 0; c < CH; ++c) s ^= acc[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_mullo(const uint32_t* in, uint64_t* out, int iters) {
  uint32_t x = in[threadIdx.x & 63] | 1u; uint32_t acc[CH];
  for (int c = 0; c < CH; ++c) acc[c] = x + c;
  BODY(asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(acc[c]) : "v"(x)));
  uint32_t s = 0; for (int c = 0; c < CH; ++c) s ^= acc[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_mulhi(const uint32_t* in, uint64_t* out, int iters) {
  uint32_t x = in[threadIdx.x & 63] | 1u; uint32_t acc[CH];
  for (int c = 0; c < CH; ++c) acc[c] = x + c;
  BODY(asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(acc[c]) : "v"(x)));
  uint32_t s = 0; for (int c = 0; c < CH; ++c) s ^= acc[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_u24(const uint32_t* in, uint64_t* out, int iters) {
  uint32_t x = in[threadIdx.x & 63] | 1u; uint32_t acc[CH];
  for (int c = 0; c < CH; ++c) acc[c] = x + c;
  BODY(asm volatile("v_mad_u32_u24 %0, %0, %1, %0" : "+v"(acc[c]) : "v"(x)));
  uint32_t s = 0; for (int c = 0; c < CH; ++c) s ^= acc[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_add32(const uint32_t* in, uint64_t* out, int iters) {
  uint32_t x = in[threadIdx.x & 63] | 1u; uint32_t acc[CH];
  for (int c = 0; c < CH; ++c) acc[c] = x + c;
  BODY(asm volatile("v_add_co_u32 %0, vcc, %0, %1" : "+v"(acc[c]) : "v"(x) : "vcc"));
  uint32_t s = 0; for (int c = 0; c < CH; ++c) s ^= acc[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_dpp(const uint32_t* in, uint64_t* out, int iters) {
  uint32_t x = in[threadIdx.x & 63] | 1u; uint32_t acc[CH];
  for (int c = 0; c < CH; ++c) acc[c] = x + c;
  BODY(asm volatile("v_mov_b32_dpp %0, %0 row_newbcast:1 row_mask:0xf bank_mask:0xf" : "+v"(acc[c])));
  uint32_t s = 0; for (int c = 0; c < CH; ++c) s ^= acc[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_mix(const uint32_t* in, uint64_t* out, int iters) {
  // one mad + one 64-bit add per step: the lazy-carry inner loop's ratio is lower than this
  uint32_t x = in[threadIdx.x & 63] | 1u; uint64_t y = x; uint64_t acc[CH], acc2[CH];
  for (int c = 0; c < CH; ++c) { acc[c] = x + c; acc2[c] = c; }
  BODY(asm volatile("v_mad_u64_u32 %0, s[40:41], %2, %2, %0\n v_lshl_add_u64 %1, %1, 0, %3" : "+v"(acc[c]), "+v"(acc2[c]) : "v"(x), "v"(y) : "s40", "s41"));
  uint64_t s = 0; for (int c = 0; c < CH; ++c) s ^= acc[c] ^ acc2[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}
#define CHECK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1); } } while (0)
typedef void (*kern_t)(const uint32_t*, uint64_t*, int);
static float timeit(kern_t k, const uint32_t* in, uint64_t* out, int blocks, int iters) {
  hipEvent_t a, b; CHECK(hipEventCreate(&a)); CHECK(hipEventCreate(&b));
  hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, 0, in, out, 1); CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(a)); hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, 0, in, out, iters);
  CHECK(hipEventRecord(b)); CHECK(hipEventSynchronize(b));
  float ms; CHECK(hipEventElapsedTime(&ms, a, b)); return ms;
}
int main() {
  hipDeviceProp_t p; CHECK(hipGetDeviceProperties(&p, 0));
  printf("device %s CUs %d clock %d kHz\n", p.gcnArchName, p.multiProcessorCount, p.clockRate);
  const int blocks = p.multiProcessorCount * 16, iters = 1000;
  uint32_t* in; uint64_t* out; CHECK(hipMalloc(&in, 4096)); CHECK(hipMalloc(&out, (size_t)blocks * 256 * 8));
  CHECK(hipMemset(in, 7, 4096));
  double ops = (double)blocks * 256 * iters * 16 * CH;
  struct { const char* name; kern_t k; } ks[] = {{"v_mad_u64_u32", k_mad}, {"v_fma_f64", k_fma}, {"v_lshl_add_u64", k_add64},
    {"v_mul_lo_u32", k_mullo}, {"v_mul_hi_u32", k_mulhi}, {"v_mad_u32_u24", k_u24}, {"v_add_co_u32", k_add32},
    {"dpp row_newbcast", k_dpp}, {"mad+add64 pair", k_mix}};
  for (auto& e : ks) {
    float ms = timeit(e.k, in, out, blocks, iters);
    printf("%-18s %8.3f ms  %7.2f T lane-instr/s  (%.2f per CU per clk)\n", e.name, ms, ops / ms / 1e9,
           ops / (ms * 1e-3) / p.multiProcessorCount / (p.clockRate * 1e3));
  }
  return 0;
}
