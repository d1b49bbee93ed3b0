// valu_probe.hip -- issue rate of the VALU instructions the kernels use.
//
// Each kernel runs 32 independent copies of one instruction per loop
// iteration (eight accumulators, unrolled four times, written as inline asm
// so that the instruction is exactly the one named), on 8 blocks of four
// wavefronts per CU.  Printed per instruction, one JSON object per line:
// wavefront instructions per second, lane operations per second (x 64) and
// SIMD cycles per wavefront instruction at the device clock.  The alignment
// and sketch rooflines divide by the device's VALU peak; this is the
// measurement that peak is checked against.
//
//   ngia2_amd/bin/ngia2_valu_probe [iterations]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {

#define CHECK(x)                                                                   \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                 \
      std::exit(1);                                                                \
    }                                                                              \
  } while (0)

enum Op {
  kXor, kAdd, kShr, kMin, kMin3, kLshlAdd, kBitop3, kPerm, kAddCo, kAddcCo, kCmp, kMulLo,
  kCndmask, kMovSdwa, kAddCoVcc, kCmpVcc, kLshlAdd64, kCmpEq64, kCmpLt64, kOps
};
const char* kNames[kOps] = {"v_xor_b32",      "v_add_u32",     "v_lshrrev_b32", "v_min_u32",
                            "v_min3_u32",     "v_lshl_add_u32", "v_bitop3_b32",  "v_perm_b32",
                            "v_add_co_u32",   "v_addc_co_u32", "v_cmp_eq_u32",  "v_mul_lo_u32",
                            "v_cndmask_b32 (e64, SGPR mask)", "v_mov_b32_sdwa (byte 1 of dst)",
                            "v_add_co_u32_e32 (carry to VCC)", "v_cmp_eq_u32_e32 (to VCC)",
                            "v_lshl_add_u64", "v_cmp_eq_u64", "v_cmp_lt_u64"};

template <int OP>
__device__ __forceinline__ void step(uint32_t& a, uint64_t& m, uint32_t b, uint32_t c) {
  if constexpr (OP == kXor) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a) : "v"(b));
  if constexpr (OP == kAdd) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a) : "v"(b));
  if constexpr (OP == kShr) asm volatile("v_lshrrev_b32 %0, 3, %0" : "+v"(a));
  if constexpr (OP == kMin) asm volatile("v_min_u32 %0, %0, %1" : "+v"(a) : "v"(b));
  if constexpr (OP == kMin3) asm volatile("v_min3_u32 %0, %0, %1, %2" : "+v"(a) : "v"(b), "v"(c));
  if constexpr (OP == kLshlAdd) asm volatile("v_lshl_add_u32 %0, %0, 3, %1" : "+v"(a) : "v"(b));
  if constexpr (OP == kBitop3)
    asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0xf4" : "+v"(a) : "v"(b), "v"(c));
  if constexpr (OP == kPerm) asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(a) : "v"(b), "v"(c));
  if constexpr (OP == kAddCo)
    asm volatile("v_add_co_u32 %0, %1, %0, %2" : "+v"(a), "=s"(m) : "v"(b));
  if constexpr (OP == kAddcCo)
    asm volatile("v_addc_co_u32 %0, %1, %0, %2, %1" : "+v"(a), "+s"(m) : "v"(b));
  if constexpr (OP == kCmp) asm volatile("v_cmp_eq_u32 %0, %1, %2" : "=s"(m) : "v"(a), "v"(b));
  if constexpr (OP == kMulLo) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a) : "v"(b));
  if constexpr (OP == kCndmask)
    asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a) : "v"(b), "s"(m));
  if constexpr (OP == kAddCoVcc) asm volatile("v_add_co_u32_e32 %0, vcc, %0, %1" : "+v"(a) : "v"(b) : "vcc");
  if constexpr (OP == kCmpVcc) asm volatile("v_cmp_eq_u32_e32 vcc, %0, %1" : : "v"(a), "v"(b) : "vcc");
  if constexpr (OP == kMovSdwa)
    asm volatile("v_mov_b32_sdwa %0, %1 dst_sel:BYTE_1 dst_unused:UNUSED_PRESERVE src0_sel:BYTE_2"
                 : "+v"(a) : "v"(b));
  // 64-bit forms: m is the accumulator (a VGPR pair) or the compare's mask
  const uint64_t bb = ((uint64_t)b << 32) | c, cc = ((uint64_t)c << 32) | b;
  if constexpr (OP == kLshlAdd64) asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(m) : "v"(bb));
  if constexpr (OP == kCmpEq64) asm volatile("v_cmp_eq_u64 %0, %1, %2" : "=s"(m) : "v"(bb), "v"(cc));
  if constexpr (OP == kCmpLt64) asm volatile("v_cmp_lt_u64 %0, %1, %2" : "=s"(m) : "v"(bb), "v"(cc));
}

template <int OP>
__global__ __launch_bounds__(256) void k_probe(uint32_t* out, uint32_t iters, uint32_t seed) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  uint32_t a[8];
  uint64_t m[8];
  #pragma unroll
  for (int u = 0; u < 8; u++) {
    a[u] = t * (2u * u + 1u) ^ seed;
    m[u] = 0;
  }
  const uint32_t b = seed * 3u + t, c = 0x0C0D0E0Fu ^ seed;
  for (uint32_t i = 0; i < iters; i++) {
    #pragma unroll
    for (int r = 0; r < 4; r++) {
      #pragma unroll
      for (int u = 0; u < 8; u++) step<OP>(a[u], m[u], b, c);
    }
  }
  uint32_t x = 0;
  #pragma unroll
  for (int u = 0; u < 8; u++) x ^= a[u] ^ (uint32_t)m[u] ^ (uint32_t)(m[u] >> 32);
  out[t] = x;
}

template <int OP>
void run(uint32_t* out, uint32_t blocks, uint32_t iters, int clock_khz, int simds) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_probe<OP>, dim3(blocks), dim3(256), 0, 0, out, 4u, 1u);  // warm-up
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(e0, 0));
  hipLaunchKernelGGL(k_probe<OP>, dim3(blocks), dim3(256), 0, 0, out, iters, 7u);
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(e1, 0));
  CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  const double waves = (double)blocks * 4.0;
  const double instr = waves * iters * 32.0;
  const double per_s = instr / (ms * 1e-3);
  const double cycles = (double)simds * (clock_khz * 1e3) / per_s;
  std::printf("{\"op\": \"%s\", \"ms\": %.4f, \"wave_instr\": %.0f, \"wave_instr_per_s\": %.4e, "
              "\"lane_ops_t_per_s\": %.2f, \"simd_cycles_per_wave_instr\": %.3f}\n",
              kNames[OP], ms, instr, per_s, per_s * 64.0 / 1e12, cycles);
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
}

}  // namespace

int main(int argc, char** argv) {
  const uint32_t iters = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 4000u;
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const uint32_t blocks = (uint32_t)prop.multiProcessorCount * 8;
  uint32_t* out = nullptr;
  CHECK(hipMalloc(&out, (size_t)blocks * 256 * 4));
  std::printf("{\"device\": \"%s\", \"cu\": %d, \"clock_khz\": %d, \"blocks\": %u, "
              "\"iterations\": %u}\n",
              prop.gcnArchName, prop.multiProcessorCount, prop.clockRate, blocks, iters);
  const int simds = prop.multiProcessorCount * 4;
  run<kXor>(out, blocks, iters, prop.clockRate, simds);
  run<kAdd>(out, blocks, iters, prop.clockRate, simds);
  run<kShr>(out, blocks, iters, prop.clockRate, simds);
  run<kMin>(out, blocks, iters, prop.clockRate, simds);
  run<kMin3>(out, blocks, iters, prop.clockRate, simds);
  run<kLshlAdd>(out, blocks, iters, prop.clockRate, simds);
  run<kBitop3>(out, blocks, iters, prop.clockRate, simds);
  run<kPerm>(out, blocks, iters, prop.clockRate, simds);
  run<kAddCo>(out, blocks, iters, prop.clockRate, simds);
  run<kAddcCo>(out, blocks, iters, prop.clockRate, simds);
  run<kCmp>(out, blocks, iters, prop.clockRate, simds);
  run<kMulLo>(out, blocks, iters, prop.clockRate, simds);
  run<kCndmask>(out, blocks, iters, prop.clockRate, simds);
  run<kMovSdwa>(out, blocks, iters, prop.clockRate, simds);
  run<kAddCoVcc>(out, blocks, iters, prop.clockRate, simds);
  run<kCmpVcc>(out, blocks, iters, prop.clockRate, simds);
  run<kLshlAdd64>(out, blocks, iters, prop.clockRate, simds);
  run<kCmpEq64>(out, blocks, iters, prop.clockRate, simds);
  run<kCmpLt64>(out, blocks, iters, prop.clockRate, simds);
  CHECK(hipFree(out));
  return 0;
}
