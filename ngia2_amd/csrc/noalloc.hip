// noalloc.hip -- ngia_align_noalloc: the kernel_dynamic contract on
// caller-owned device memory, for callers that cannot allocate (code under
// HIP graph capture, serving loops that forbid hipMalloc).  The only device
// memory beyond the caller's inputs and outputs is the plain kernel's LCS row
// scratch, and it lives in the file-scope array below: one copy per GPU, made
// when the code object loads and held until the process ends.
#include <mutex>

#include "capi_internal.hpp"

using namespace ngia;

namespace {

constexpr uint64_t kScratchWords = NGIA_NOALLOC_SCRATCH_BYTES / 4;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
uint32_t* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU

}  // namespace

// The row scratch.  In host code this name is a host-side stand-in, not a
// device address: host code uses only the address hipGetSymbolAddress gives
// for the current device (noalloc_setup).
static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace ngia {

// Called by ngia_open with `device` current: the device address of this GPU's
// copy of the scratch, looked up once per device.
uint32_t* noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) throw HipError{NGIA_ERR_ARG, "device index above 63"};
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    NGIA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError{NGIA_ERR_HIP, "hipGetSymbolAddress(scratch) gave a null address"};
    g_scratch[device] = (uint32_t*)p;
  }
  return g_scratch[device];
}

}  // namespace ngia

extern "C" int ngia_align_noalloc(ngia_ctx* ctx, const uint32_t* packed, const uint64_t* word_off,
                                  uint32_t* jobs, uint32_t count, float threshold,
                                  uint32_t max_len, void* stream) {
  if (!ctx) return fail(nullptr, NGIA_ERR_ARG, "null context");
  if (count && (!packed || !word_off || !jobs)) return fail(ctx, NGIA_ERR_ARG, "null argument");
  if (max_len > NGIA_MAX_LEN)
    return fail(ctx, NGIA_ERR_ARG, "max_len above 65534 (the longest record makeDB keeps)");
  if (count == 0) return NGIA_OK;
  return guarded(ctx, [&] {
    if (!ctx->noalloc_scratch) throw HipError{NGIA_ERR_STATE, "context has no scratch address"};
    AlignArgs a;
    a.packed = packed;
    a.word_off = word_off;
    a.codes = nullptr;  // target codes come from the planes
    a.code_off = nullptr;
    a.thr = threshold;
    a.raw_lcs = ctx->raw_lcs ? 1u : 0u;
    const uint32_t nw_cap = std::max<uint32_t>(1, (max_len + 31) / 32);
    launch_align_plain_tiers(a, jobs, count, ctx->noalloc_scratch, kScratchWords, nw_cap,
                             stream ? (hipStream_t)stream : ctx->stream);
    return NGIA_OK;
  });
}
