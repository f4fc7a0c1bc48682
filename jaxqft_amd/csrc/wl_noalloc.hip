// wl_apply without device allocation: the operator's three FP64 staging
// buffers live in a statically sized scratch array that is part of the code
// object, one copy per device (each GPU that loads the library holds it until
// the process ends).  Host code never uses the array's name as an address:
// its device address is taken once per device, with that device current, at
// context creation (wl_ctx_create calls noalloc_init_device), and kept per
// device.  The staging buffers need no clearing: the pack kernel writes every
// site of every column before the stencil reads it, and the padding sites of
// a plane are never read.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/wilson_amd.h"

namespace wl {
int report_error(int code, const char* msg);
size_t apply_buffer_bytes(const wl_ctx* c, int nrhs);
int ctx_device(const wl_ctx* c);
int apply_checked(wl_ctx* c, wl_gauge* g, int op, const void* in, void* out, int nrhs, int dtype, void* b0, void* b1,
                  void* b2, hipStream_t st);
}  // namespace wl

static __device__ __align__(256) unsigned char scratch[WL_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};
hipError_t g_err[MAX_DEVICES] = {};
bool g_done[MAX_DEVICES] = {};
}  // namespace

namespace wl {

// Takes the device address of `scratch` on device dev (the current device)
// the first time a context is created there.
void noalloc_init_device(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_done[dev]) return;
  g_done[dev] = true;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_err[dev] = e;
  if (e == hipSuccess) {
    g_ptr[dev] = ptr;
  } else {
    (void)hipGetLastError();
    g_ptr[dev] = nullptr;
  }
}

}  // namespace wl

static inline size_t round256(size_t b) { return (b + 255) / 256 * 256; }

extern "C" int wl_apply_noalloc(wl_ctx* c, wl_gauge* g, int op, const void* in, void* out, int nrhs, int dtype,
                                void* stream) {
  using namespace wl;
  if (!c) return report_error(WL_ERR_BAD_ARG, "null context");
  if (nrhs < 1) return report_error(WL_ERR_BAD_ARG, "nrhs must be >= 1");
  const int dev = ctx_device(c);
  void* base = nullptr;
  hipError_t why = hipSuccess;
  if (dev >= 0 && dev < MAX_DEVICES) {
    std::lock_guard<std::mutex> lk(g_mu);
    base = g_ptr[dev];
    why = g_err[dev];
  }
  if (!base) {
    char msg[256];
    snprintf(msg, sizeof(msg), "wl_apply_noalloc: no scratch address on device %d (%s)", dev,
             why == hipSuccess ? "not initialised" : hipGetErrorString(why));
    return report_error(WL_ERR_HIP, msg);
  }
  const size_t per = round256(apply_buffer_bytes(c, nrhs));
  const int nbuf = op == WL_OP_DDAGD ? 3 : 2;
  if (per > WL_NOALLOC_SCRATCH_BYTES / nbuf) {
    char msg[256];
    snprintf(msg, sizeof(msg),
             "wl_apply_noalloc: %d x %zu bytes of staging for %d right-hand sides exceed the %zu-byte scratch "
             "array (use wl_apply)", nbuf, per, nrhs, (size_t)WL_NOALLOC_SCRATCH_BYTES);
    return report_error(WL_ERR_BAD_ARG, msg);
  }
  char* b = static_cast<char*>(base);
  return apply_checked(c, g, op, in, out, nrhs, dtype, b, b + per, nbuf == 3 ? b + 2 * per : nullptr,
                       (hipStream_t)stream);
}
