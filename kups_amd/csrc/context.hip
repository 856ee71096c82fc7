// Context, scratch and error plumbing of the C ABI.
#include <string>
#include "common.hpp"

namespace kupsmi {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int scratch(kupsmi_ctx* ctx, int slot, size_t bytes, void** out) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
  KUPSMI_REQUIRE(slot >= 0 && slot < 16, "bad scratch slot");
  Scratch& s = ctx->slots[slot];
  if (s.bytes < bytes) {
    KUPSMI_REQUIRE(!ctx->frozen,
                   "scratch of a frozen context cannot grow (recorded work uses its buffers): "
                   "size it before freezing, or use another context");
    // Grow geometrically so a slowly growing workload reallocates rarely.
    size_t nb = s.bytes * 2 > bytes ? s.bytes * 2 : bytes;
    nb = (nb + 255) & ~(size_t)255;
    if (s.ptr) KUPSMI_HIP(hipFree(s.ptr));
    s.ptr = nullptr;
    s.bytes = 0;
    KUPSMI_HIP(hipMalloc(&s.ptr, nb));
    s.bytes = nb;
  }
  *out = s.ptr;
  return KUPSMI_OK;
}

}  // namespace kupsmi

using namespace kupsmi;

extern "C" int kupsmi_version(void) { return 100; }

extern "C" const char* kupsmi_last_error(void) { return g_last_error.c_str(); }

extern "C" int kupsmi_ctx_create(kupsmi_ctx** out) {
  KUPSMI_REQUIRE(out != nullptr, "null output pointer");
  kupsmi_ctx* c = new (std::nothrow) kupsmi_ctx();
  KUPSMI_REQUIRE(c != nullptr, "out of host memory");
  KUPSMI_HIP(hipGetDevice(&c->device));
  const hipError_t e = hipEventCreateWithFlags(&c->phase, hipEventDisableTiming);
  if (e != hipSuccess) {
    delete c;
    KUPSMI_HIP(e);
  }
  *out = c;
  return KUPSMI_OK;
}

extern "C" int kupsmi_ctx_record_pairs(kupsmi_ctx* ctx, int64_t capacity, int32_t* ij,
                                      int32_t* shift, uint8_t* terms, unsigned long long* count) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
  KUPSMI_REQUIRE(capacity >= 0 && count != nullptr, "count pointer and capacity >= 0 required");
  KUPSMI_REQUIRE(capacity == 0 || (ij && shift), "ij and shift buffers required");
  ctx->rec = PairRecord{capacity, ij, shift, terms, count};
  ctx->rec_armed = 1;
  return KUPSMI_OK;
}

extern "C" int kupsmi_ctx_freeze(kupsmi_ctx* ctx, int frozen) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
  ctx->frozen = frozen != 0;
  return KUPSMI_OK;
}

extern "C" int kupsmi_ctx_destroy(kupsmi_ctx* ctx) {
  if (!ctx) return KUPSMI_OK;
  int prev = 0;
  KUPSMI_HIP(hipGetDevice(&prev));
  KUPSMI_HIP(hipSetDevice(ctx->device));
  for (auto& s : ctx->slots)
    if (s.ptr) (void)hipFree(s.ptr);
  if (ctx->phase) (void)hipEventDestroy(ctx->phase);
  (void)hipSetDevice(prev);
  delete ctx;
  return KUPSMI_OK;
}
