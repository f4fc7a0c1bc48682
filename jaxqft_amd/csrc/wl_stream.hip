// HBM streaming probes: read-only, write-only and copy kernels of 16 bytes
// per lane, the reference rates against which the stencils' bandwidth is
// read.  A workgroup of 256 lanes moves 4 KiB per step; UNROLL steps are
// issued back to back before their data is used, and the grid either covers
// the buffer once (blocks = 0) or strides over it (a fixed number of
// workgroups).  NT_HINT selects non-temporal loads and stores.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/wilson_amd.h"
#include "wl_kernels.h"

namespace wl {
int report_error(int code, const char* msg);

template <bool NT_HINT>
__device__ __forceinline__ uint4 ld16(const uint4* p) {
  if constexpr (NT_HINT) {
    uint4 v;
    v.x = __builtin_nontemporal_load(&p->x);
    v.y = __builtin_nontemporal_load(&p->y);
    v.z = __builtin_nontemporal_load(&p->z);
    v.w = __builtin_nontemporal_load(&p->w);
    return v;
  } else {
    return *p;
  }
}

template <bool NT_HINT>
__device__ __forceinline__ void st16(uint4* p, uint4 v) {
  if constexpr (NT_HINT) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
    __builtin_nontemporal_store(v.z, &p->z);
    __builtin_nontemporal_store(v.w, &p->w);
  } else {
    *p = v;
  }
}

// KIND 0: read (XOR of everything read, stored only if it equals a value no
// buffer of this probe holds, so the loads stay live); 1: write; 2: copy.
template <int KIND, bool NT_HINT, int UNROLL>
__global__ __launch_bounds__(NT) void k_stream(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n,
                                                unsigned* sink) {
  const size_t stride = (size_t)gridDim.x * NT;
  size_t i = (size_t)blockIdx.x * NT + threadIdx.x;
  uint32_t acc = 0;
  const uint4 val = make_uint4(0x3f800000u, 0u, 0x3f800000u, 0u);
  for (; i + (UNROLL - 1) * stride < n; i += UNROLL * stride) {
    uint4 v[UNROLL];
    if (KIND != 1) {
#pragma unroll
      for (int u = 0; u < UNROLL; ++u) v[u] = ld16<NT_HINT>(src + i + u * stride);
    }
#pragma unroll
    for (int u = 0; u < UNROLL; ++u) {
      if (KIND == 0) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
      if (KIND == 1) st16<NT_HINT>(dst + i + u * stride, val);
      if (KIND == 2) st16<NT_HINT>(dst + i + u * stride, v[u]);
    }
  }
  for (; i < n; i += stride) {
    if (KIND == 0) {
      const uint4 v = ld16<NT_HINT>(src + i);
      acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    if (KIND == 1) st16<NT_HINT>(dst + i, val);
    if (KIND == 2) st16<NT_HINT>(dst + i, ld16<NT_HINT>(src + i));
  }
  if (KIND == 0 && acc == 0x9e3779b9u) sink[0] = acc;
}

template <int KIND, bool NT_HINT>
static hipError_t launch_kind(int unroll, const uint4* src, uint4* dst, size_t n, unsigned blocks, unsigned* sink,
                              hipStream_t st) {
  switch (unroll) {
    case 1: hipLaunchKernelGGL((k_stream<KIND, NT_HINT, 1>), dim3(blocks), dim3(NT), 0, st, src, dst, n, sink); break;
    case 2: hipLaunchKernelGGL((k_stream<KIND, NT_HINT, 2>), dim3(blocks), dim3(NT), 0, st, src, dst, n, sink); break;
    case 4: hipLaunchKernelGGL((k_stream<KIND, NT_HINT, 4>), dim3(blocks), dim3(NT), 0, st, src, dst, n, sink); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace wl

using namespace wl;

extern "C" int wl_stream_probe(int kind, int nontemporal, int unroll, const void* src, void* dst, size_t bytes,
                               int blocks, void* stream) {
  if (kind < WL_STREAM_READ || kind > WL_STREAM_COPY) return report_error(WL_ERR_BAD_ARG, "stream probe: bad kind");
  if (unroll != 1 && unroll != 2 && unroll != 4) return report_error(WL_ERR_BAD_ARG, "stream probe: unroll 1, 2 or 4");
  if (blocks < 0 || bytes == 0 || bytes % 16 != 0)
    return report_error(WL_ERR_BAD_ARG, "stream probe: bytes must be a positive multiple of 16, blocks >= 0");
  if ((kind != WL_STREAM_WRITE && (!src || (uintptr_t)src % 16)) || !dst || (uintptr_t)dst % 16)
    return report_error(WL_ERR_BAD_ARG, "stream probe: null or unaligned buffer");
  const size_t n = bytes / 16;
  size_t nb = blocks > 0 ? (size_t)blocks : (n + (size_t)NT * unroll - 1) / ((size_t)NT * unroll);
  if (nb > 0x7fffffffULL) return report_error(WL_ERR_BAD_ARG, "stream probe: too many workgroups");
  const uint4* s = reinterpret_cast<const uint4*>(src);
  uint4* d = reinterpret_cast<uint4*>(dst);
  // the read kernel's sink is the first word of dst, written only by a lane
  // whose XOR of what it read equals the sentinel
  unsigned* sink = reinterpret_cast<unsigned*>(dst);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  const unsigned g = (unsigned)nb;
  if (kind == WL_STREAM_READ)
    e = nontemporal ? launch_kind<0, true>(unroll, s, d, n, g, sink, st) : launch_kind<0, false>(unroll, s, d, n, g, sink, st);
  else if (kind == WL_STREAM_WRITE)
    e = nontemporal ? launch_kind<1, true>(unroll, s, d, n, g, sink, st) : launch_kind<1, false>(unroll, s, d, n, g, sink, st);
  else
    e = nontemporal ? launch_kind<2, true>(unroll, s, d, n, g, sink, st) : launch_kind<2, false>(unroll, s, d, n, g, sink, st);
  return e == hipSuccess ? WL_OK : report_error(WL_ERR_HIP, hipGetErrorString(e));
}
