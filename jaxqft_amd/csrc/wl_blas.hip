// Fused BLAS-1 kernels of the CG iteration, one site per thread, all spinor
// planes of the site per thread (so the FP16 per-site scale is local), FP64
// accumulation of every norm and dot product.
#include "wl_kernels.h"
#include "wl_reduce.h"

namespace wl {



// Every vector a BLAS kernel touches is read or written once: streaming
// (non-temporal) accesses (measured at 32^4 on 12 FP16 columns: p' = r + beta
// p 378 -> 352 us, and 319 us together with the column-slowest order below).
template <int P, int NC, class T>
__device__ __forceinline__ void load_as(const void* base, size_t V, size_t site, cx<T> (&v)[4 * NC]) {
  using A = typename PT<P>::acc;
  cx<A> t[4 * NC];
  SpinorIO<P, NC>::template load<true>(base, V, site, t);
#pragma unroll
  for (int c = 0; c < 4 * NC; ++c) v[c] = {(T)t[c].re, (T)t[c].im};
}

template <int P, int NC, class T>
__device__ __forceinline__ void store_as(void* base, size_t V, size_t site, const cx<T> (&v)[4 * NC]) {
  using A = typename PT<P>::acc;
  cx<A> t[4 * NC];
#pragma unroll
  for (int c = 0; c < 4 * NC; ++c) t[c] = {(A)v[c].re, (A)v[c].im};
  SpinorIO<P, NC>::template store<true>(base, V, site, t);
}

// A workgroup loops over its (column, tile) units: one unit per workgroup
// normally, several when a launch gated on pending reliable updates is capped
// (blas_grid), so that a mostly idle launch dispatches few workgroups; a
// capped launch first checks, with one round of loads, whether the gate
// admits any column at all.  Units are ordered column-slowest: each column is
// streamed as one contiguous run.
#define WL_BLAS_PROLOGUE                                            \
  const unsigned total_ = (unsigned)a.ntiles * (unsigned)a.ncol;    \
  if (gridDim.x < total_ && !gate_any_col(a.gate, a.red, a.ncol))   \
    return;                                                         \
  for (unsigned u_ = blockIdx.x; u_ < total_; u_ += gridDim.x) {    \
  const int col = (int)(u_ / (unsigned)a.ntiles);                   \
  const int tile = (int)(u_ % (unsigned)a.ntiles);                  \
  if (gate_skip_col(a.gate, a.red, col)) continue;                  \
  const size_t V = (size_t)a.Vp;                                    \
  const size_t site = (size_t)tile * NT + threadIdx.x;              \
  const bool valid = site < (size_t)a.V;                            \
  const size_t s = valid ? site : 0;
#define WL_BLAS_EPILOGUE }

// |a|^2 -> action
template <int P, int NC>
__global__ __launch_bounds__(NT) void k_norm2(BlasArgs a) {
  WL_BLAS_PROLOGUE
  using T = typename PT<P>::acc;
  cx<T> v[4 * NC];
  load_as<P, NC>((char*)a.a + col * a.col_bytes_a, V, s, v);
  double part = 0.0;
#pragma unroll
  for (int c = 0; c < 4 * NC; ++c) part += (double)v[c].re * v[c].re + (double)v[c].im * v[c].im;
  write_partial(a.red, col, tile, a.ntiles, valid ? part : 0.0);
  WL_BLAS_EPILOGUE
}

// p' = r + beta p.   b=r; p from a, or, when d is set, from a or d as the
// column's p_par says (buffer 0 or 1); p' goes to c (null: back to the source)
template <int P, int NC>
__global__ __launch_bounds__(NT) void k_xpay(BlasArgs a) {
  WL_BLAS_PROLOGUE
  using T = typename PT<P>::acc;
  const ColState& cs = a.red.cols[col];
  const T beta = (T)cs.beta;
  void* src = (a.d && cs.p_par) ? a.d : a.a;
  cx<T> p[4 * NC], r[4 * NC];
  load_as<P, NC>((char*)src + col * a.col_bytes_a, V, s, p);
  load_as<P, NC>((char*)a.b + col * a.col_bytes_b, V, s, r);
#pragma unroll
  for (int c = 0; c < 4 * NC; ++c) {
    p[c].re = r[c].re + beta * p[c].re;
    p[c].im = r[c].im + beta * p[c].im;
  }
  if (valid) store_as<P, NC>((char*)(a.c ? a.c : src) + col * a.col_bytes_a, V, s, p);
  WL_BLAS_EPILOGUE
}

// x += y (+ alpha_prev p when a pairwise column has alpha p pending); y = 0.
// a=x (PX), b=y (PY), c and d = the two search-direction buffers (PP; the
// column's p_par picks the one holding the pending p; null: none pending)
template <int PX, int PY, int PP, int NC>
__global__ __launch_bounds__(NT) void k_accumulate(BlasArgs a) {
  WL_BLAS_PROLOGUE
  cx<double> x[4 * NC], y[4 * NC];
  load_as<PX, NC>((char*)a.a + col * a.col_bytes_a, V, s, x);
  load_as<PY, NC>((char*)a.b + col * a.col_bytes_b, V, s, y);
  const ColState& cs = a.red.cols[col];
  if (a.c && cs.pairwise && cs.pend) {
    cx<double> p[4 * NC];
    load_as<PP, NC>((char*)(cs.p_par ? a.d : a.c) + col * a.col_bytes_c, V, s, p);
    const double ap = cs.alpha_prev;
#pragma unroll
    for (int c = 0; c < 4 * NC; ++c) {
      y[c].re += ap * p[c].re;
      y[c].im += ap * p[c].im;
    }
  }
  cx<double> z[4 * NC];
#pragma unroll
  for (int c = 0; c < 4 * NC; ++c) {
    x[c].re += y[c].re;
    x[c].im += y[c].im;
    z[c] = {0.0, 0.0};
  }
  if (valid) {
    store_as<PX, NC>((char*)a.a + col * a.col_bytes_a, V, s, x);
    store_as<PY, NC>((char*)a.b + col * a.col_bytes_b, V, s, z);
  }
  WL_BLAS_EPILOGUE
}

// out = in (precision conversion).   a=out (PO), b=in (PI)
template <int PO, int PI, int NC>
__global__ __launch_bounds__(NT) void k_convert(BlasArgs a) {
  WL_BLAS_PROLOGUE
  cx<double> v[4 * NC];
  load_as<PI, NC>((char*)a.b + col * a.col_bytes_b, V, s, v);
  if (valid) store_as<PO, NC>((char*)a.a + col * a.col_bytes_a, V, s, v);
  WL_BLAS_EPILOGUE
}

__global__ __launch_bounds__(NT) void k_reduce_finish(Reduction red, int nblk, int gate) {
  const int col = blockIdx.x;
  if (gate_skip_col(gate, red, col)) return;
  __shared__ double lds[NT / 64];
  const double* parts = red.partials + (size_t)col * nblk;
  double s = 0.0;
  for (int i = threadIdx.x; i < nblk; i += NT) s += parts[i];
  s = block_sum<NT>(s, lds);
  if (threadIdx.x == 0) {
    apply_action(red.action, red.cols[col], s, red.par);
    if (red.then_finalize) finalize_col(red.cols[col]);
  }
}

hipError_t launch_reduce_finish(const Reduction& red, int ncol, int nblk, int gate, hipStream_t stream) {
  hipLaunchKernelGGL(k_reduce_finish, dim3(ncol), dim3(NT), 0, stream, red, nblk, gate);
  return hipGetLastError();
}

// A newly admitted column: targets from |b|^2, then the next reliable update
// computes its initial residual.
__global__ void k_admit_finish(ColState* cols, int col, double tol2) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ColState& s = cols[col];
    s.tol2b2 = tol2 * s.b2;
    s.stop2 = tol2 * s.b2;
    s.need_update = 1;
  }
}

hipError_t launch_admit_finish(ColState* cols, int col, double tol2, hipStream_t stream) {
  hipLaunchKernelGGL(k_admit_finish, dim3(1), dim3(64), 0, stream, cols, col, tol2);
  return hipGetLastError();
}

// First update of a solve from a zero initial guess: the true residual is b
// itself, so its norm is |b|^2 and the convergence decision of a true-residual
// reduction is taken on that (a zero right-hand side converges here).
__global__ void k_trueres_from_b2(ColState* cols, int ncol, int then_finalize) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  ColState& s = cols[c];
  if (!(s.need_update && !s.done)) return;
  apply_action(RA_TRUERES, s, s.b2);
  if (then_finalize) finalize_col(s);
}

// After a reliable update: restart the iterated residual from the true one.
__global__ void k_finalize_update(ColState* cols, int ncol) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  finalize_col(cols[c]);
}

// ---------------------------------------------------------------------------
// host dispatch
// ---------------------------------------------------------------------------
#define WL_DISPATCH_NC(KERNEL, ...)                                                \
  ((nc == 3 ? KERNEL<__VA_ARGS__, 3><<<grid, block, 0, stream>>>(a)                 \
            : KERNEL<__VA_ARGS__, 1><<<grid, block, 0, stream>>>(a)),               \
   hipGetLastError())

static inline dim3 blas_grid(const BlasArgs& a) {
  const unsigned units = (unsigned)(a.ntiles * a.ncol);
  // gated on pending reliable updates: about the resident workgroups (8 per CU)
  return dim3((a.gate == 2 && units > 2048u) ? 2048u : units);
}

static hipError_t norm2_main(int p, int nc, BlasArgs a, hipStream_t stream) {
  dim3 grid = blas_grid(a), block(NT);
  if (p == F64) return WL_DISPATCH_NC(k_norm2, F64);
  if (p == F32) return WL_DISPATCH_NC(k_norm2, F32);
  return WL_DISPATCH_NC(k_norm2, F16);
}

hipError_t launch_norm2(int p, int nc, BlasArgs a, hipStream_t stream) {
  hipError_t e = norm2_main(p, nc, a, stream);
  if (e != hipSuccess) return e;
  return launch_reduce_finish(a.red, a.ncol, a.ntiles, a.gate, stream);
}

hipError_t launch_xpay(int p, int nc, BlasArgs a, hipStream_t stream) {
  dim3 grid = blas_grid(a), block(NT);
  if (p == F64) return WL_DISPATCH_NC(k_xpay, F64);
  if (p == F32) return WL_DISPATCH_NC(k_xpay, F32);
  return WL_DISPATCH_NC(k_xpay, F16);
}

// pp: precision of the pending search direction a.c (inner precision)
hipError_t launch_accumulate(int px, int py, int pp, int nc, BlasArgs a, hipStream_t stream) {
  dim3 grid = blas_grid(a), block(NT);
  if (px != F64) return hipErrorInvalidValue;
  if (pp == F32) {
    if (py == F32) return WL_DISPATCH_NC(k_accumulate, F64, F32, F32);
    if (py == F64) return WL_DISPATCH_NC(k_accumulate, F64, F64, F32);
    if (py == F16) return WL_DISPATCH_NC(k_accumulate, F64, F16, F32);
  } else if (pp == F16) {
    if (py == F32) return WL_DISPATCH_NC(k_accumulate, F64, F32, F16);
    if (py == F64) return WL_DISPATCH_NC(k_accumulate, F64, F64, F16);
    if (py == F16) return WL_DISPATCH_NC(k_accumulate, F64, F16, F16);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_convert(int po, int pi, int nc, BlasArgs a, hipStream_t stream) {
  dim3 grid = blas_grid(a), block(NT);
  if (po == F32 && pi == F64) return WL_DISPATCH_NC(k_convert, F32, F64);
  if (po == F16 && pi == F64) return WL_DISPATCH_NC(k_convert, F16, F64);
  if (po == F64 && pi == F64) return WL_DISPATCH_NC(k_convert, F64, F64);
  if (po == F64 && pi == F32) return WL_DISPATCH_NC(k_convert, F64, F32);
  if (po == F64 && pi == F16) return WL_DISPATCH_NC(k_convert, F64, F16);
  if (po == F32 && pi == F32) return WL_DISPATCH_NC(k_convert, F32, F32);
  return hipErrorInvalidValue;
}

hipError_t launch_trueres_from_b2(ColState* cols, int ncol, int then_finalize, hipStream_t stream) {
  hipLaunchKernelGGL(k_trueres_from_b2, dim3((ncol + 63) / 64), dim3(64), 0, stream, cols, ncol, then_finalize);
  return hipGetLastError();
}

hipError_t launch_finalize_update(ColState* cols, int ncol, hipStream_t stream) {
  hipLaunchKernelGGL(k_finalize_update, dim3((ncol + 63) / 64), dim3(64), 0, stream, cols, ncol);
  return hipGetLastError();
}

}  // namespace wl
