// Host side of libwilson_amd: contexts, resident gauge fields, internal
// fields, the operator entry point and the normal-equation CG engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wilson_amd.h"
#include "wl_device.h"
#include "wl_kernels.h"

namespace wl {
hipError_t launch_dslash(int prec, int nc, int recon, int op, int epi, const DslashArgs& a, hipStream_t stream);
hipError_t launch_norm2(int p, int nc, BlasArgs a, hipStream_t stream);
hipError_t launch_xpay(int p, int nc, BlasArgs a, hipStream_t stream);
hipError_t launch_accumulate(int px, int py, int pp, int nc, BlasArgs a, hipStream_t stream);
hipError_t launch_convert(int po, int pi, int nc, BlasArgs a, hipStream_t stream);
hipError_t launch_finalize_update(ColState* cols, int ncol, hipStream_t stream);
hipError_t launch_trueres_from_b2(ColState* cols, int ncol, int then_finalize, hipStream_t stream);
hipError_t launch_admit_finish(ColState* cols, int col, double tol2, hipStream_t stream);
hipError_t pack_spinor(int prec, int nc, const void* src, int dtype, void* dst, long long V, long long Vp, int ncol,
                       size_t col_bytes, hipStream_t st);
hipError_t unpack_spinor(int prec, int nc, const void* src, void* dst, int dtype, long long V, long long Vp, int ncol,
                         size_t col_bytes, hipStream_t st);
hipError_t pack_gauge_f64(int nc, const void* src, int dtype, int layout, void* dst, long long V, long long Vp,
                          int ncfg, size_t cfg_bytes, size_t dir_bytes, hipStream_t st);
hipError_t row2_deviation(const void* src, long long V, long long Vp, int ncfg, size_t cfg_bytes, size_t dir_bytes,
                          unsigned long long* out, hipStream_t st);
hipError_t convert_gauge(int prec, int nc, int recon, const void* src, void* dst, long long V, long long Vp, int ncfg,
                         size_t src_cfg_bytes, size_t src_dir_bytes, size_t cfg_bytes, size_t dir_bytes,
                         hipStream_t st);
hipError_t launch_eo(int prec, int nc, int recon, int dag, int mode, int epi, const DslashArgs& a, hipStream_t stream);
hipError_t launch_pf_force(int nc, int dtype, const ForceArgs& a, int ncfg, hipStream_t st);
hipError_t launch_lime_decode(int word_bytes, int out_double, const LimeArgs& a, size_t lds_bytes, hipStream_t st);
void noalloc_init_device(int dev);
hipError_t pack_spinor_eo(int prec, int nc, const void* src, int dtype, void* dst, const Geom& g, long long Vh,
                          long long Vhp, int q, int ncol, size_t col_bytes, hipStream_t st);
hipError_t unpack_spinor_eo(int prec, int nc, const void* src, void* dst, int dtype, const Geom& g, long long Vhp, int q,
                            int ncol, size_t col_bytes, hipStream_t st);
hipError_t convert_gauge_eo(int prec, int nc, int recon, const void* src, void* dst, const Geom& g, long long Vh,
                            long long Vhp, int ncfg, size_t src_cfg_bytes, size_t src_dir_bytes, size_t cfg_bytes,
                            size_t dir_bytes, hipStream_t st);
}  // namespace wl

using namespace wl;

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace wl {
// Records the message of a failure found in another translation unit (the
// calling thread's wl_last_error) and returns code.
int report_error(int code, const char* msg) { return fail(code, "%s", msg); }
}  // namespace wl

#define HIPCHK(expr)                                                                            \
  do {                                                                                          \
    hipError_t _e = (expr);                                                                     \
    if (_e != hipSuccess) return fail(WL_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
  } while (0)

#define CHK(expr)            \
  do {                       \
    int _r = (expr);         \
    if (_r != WL_OK) return _r; \
  } while (0)

// Device buffer.  Each allocation starts at a staggered offset (a multiple
// of 9 cache lines) so that the same site of different fields does not map
// to the same L2 set.
static std::atomic<unsigned> g_alloc_color{0};
struct DevBuf {
  void* p = nullptr;    // usable base
  void* raw = nullptr;  // allocation
  size_t n = 0;
  int ensure(size_t bytes) {
    if (bytes <= n) return WL_OK;
    release();
    const size_t off = (size_t)(g_alloc_color++ % 8) * 1152;
    hipError_t e = hipMalloc(&raw, bytes + 8 * 1152);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      raw = nullptr;
      return fail(e == hipErrorOutOfMemory ? WL_ERR_OOM : WL_ERR_HIP, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    p = static_cast<char*>(raw) + off;
    n = bytes;
    return WL_OK;
  }
  void release() {
    if (raw) (void)hipFree(raw);
    raw = nullptr;
    p = nullptr;
    n = 0;
  }
};

struct wl_ctx {
  wl_geom geom;
  wl_params params;
  Geom g;
  long long V;
  long long Vp;  // padded plane stride
  int device;
  std::mutex mu;
  // workspace
  DevBuf partials, cols;
  DevBuf ws[12];
  long long Vh = 0;   // sites of one parity (0: even-odd operators unavailable, some L odd)
  long long Vhp = 0;  // plane stride of compact (one-parity) fields
  ColState* host_cols = nullptr;  // pinned mirror
  size_t host_cols_n = 0;
  // stencil launch tuning per [single-RHS / multi-RHS launch][precision]
  int march[2][3] = {{0, 0, 0}, {0, 0, 0}};      // x0 sites per thread (0 = automatic)
  int mrhs[3] = {0, 0, 0};                       // right-hand sides per thread (0 = automatic)
  int xcd_remap[2][3] = {{1, 1, 1}, {1, 1, 1}};  // XCD-contiguous workgroup order
  int slab[2][3] = {{0, 0, 0}, {0, 0, 0}};       // x1 extent of a slab (0 = flat site order)
  int slab2[2][3] = {{0, 0, 0}, {0, 0, 0}};      // x2 extent of a slab (0 = whole x1 planes)
  int cols_outer[2][3] = {{0, 0, 0}, {0, 0, 0}}; // column order (0 fastest, n >= 1 blocks of n columns or groups)
  int blocks_per_cu[2][3] = {{0, 0, 0}, {0, 0, 0}};  // resident workgroup cap (0 = none)
  int block_shape[2][3] = {{0, 0, 0}, {0, 0, 0}};    // 4D box preset of one-site kernels (0 = none)
  int plane[2][3] = {{0, 0, 0}, {0, 0, 0}};          // x1 march of the LDS plane kernel (0 = off)
  int lds_box[2][3] = {{0, 0, 0}, {0, 0, 0}};        // LDS box kernel preset (0 = off; F16 SU(3))
  int eo_slab[2][3] = {{0, 0, 0}, {0, 0, 0}};        // x1 slab of the even-odd hop (0 = flat)
  int lds_per_cu = 0;     // bytes of LDS per CU
  int lds_per_block = 0;  // largest LDS allocation of one workgroup
  // stream order of calls sharing the workspace: the end of the last call,
  // recorded on the stream it ran on
  hipEvent_t order_ev = nullptr;
  hipStream_t order_st = nullptr;
  bool order_set = false;
  // kernel timing of CG iterations (wl_ctx_set_profiling)
  struct Prof {
    static constexpr int SLOTS = 64;  // iterations timed per host-driven block
    bool on = false;
    hipEvent_t ev[SLOTS][4] = {};
    int par[SLOTS] = {};
    int n = 0;
    double sum_op[2] = {0, 0}, sum_dag[2] = {0, 0}, sum_xpay = 0;
    int cnt[2] = {0, 0};
    int no_update = 0;  // iterations queued without the fused update launches
  } prof;
};

// Calls on one context share its workspace (ws, partials, cols).  A call on
// a stream other than the previous call's first waits for the end of that
// call, and every call records its own end, so calls issued from different
// streams are ordered as issued.  Under graph capture the caller orders the
// captured work itself (an event of the uncaptured stream cannot be waited
// on inside a capture): nothing is waited for or recorded there.
struct StreamOrder {
  wl_ctx* c;
  hipStream_t st;
  bool on;
  StreamOrder(wl_ctx* ctx, hipStream_t s) : c(ctx), st(s), on(true) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (!c->order_ev || (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)) {
      on = false;
      return;
    }
    (void)hipGetLastError();
    if (c->order_set && c->order_st != st) (void)hipStreamWaitEvent(st, c->order_ev, 0);
  }
  ~StreamOrder() {
    if (!on) return;
    if (hipEventRecord(c->order_ev, st) == hipSuccess) {
      c->order_st = st;
      c->order_set = true;
    } else {
      (void)hipGetLastError();
    }
  }
};

struct wl_gauge {
  wl_ctx* ctx;
  int ncfg;
  DevBuf master;  // F64, full links
  std::map<int, DevBuf> enc;  // key prec*100 + recon
  double row2_dev = 1.0;  // SU(3): largest |U_2j - conj(U_0 x U_1)_j| over the links (measured at upload)
};

// Links whose stored third row equals the one rebuilt from the first two to
// this absolute accuracy give the same FP64 operator from a 12-real encoding
// (each term of D changes by at most ~1e-14 of its size, far below the 1e-10
// true-residual targets).  Configurations read in single precision, or not
// projected to SU(3), fail the test and keep the full 18 reals.
constexpr double EXACT_RECON12_TOL = 1e-14;

// Link encoding of the FP64 operator in the reliable updates and true
// residuals: 12 reals when that reproduces the stored links, else 18.
static int update_recon(const wl_gauge* g) {
  return (g->ctx->geom.nc == 3 && g->row2_dev <= EXACT_RECON12_TOL) ? 12 : 18;
}

struct wl_field {
  wl_ctx* ctx;
  int prec;
  int ncol;
  size_t col_bytes;
  DevBuf buf;
};

static inline size_t cfg_bytes(const wl_ctx* c, int prec, int recon) {
  return 4 * link_dir_bytes(prec, c->geom.nc, recon, (size_t)c->Vp);
}
static inline size_t col_bytes_of(const wl_ctx* c, int prec) { return spinor_col_bytes(prec, c->geom.nc, (size_t)c->Vp); }
// compact one-parity fields
static inline size_t col_bytes_eo(const wl_ctx* c, int prec) { return spinor_col_bytes(prec, c->geom.nc, (size_t)c->Vhp); }
static inline size_t cfg_bytes_eo(const wl_ctx* c, int prec, int recon) {
  return 8 * link_dir_bytes(prec, c->geom.nc, recon, (size_t)c->Vhp);
}

static int check_ctx(const wl_ctx* c) {
  if (!c) return fail(WL_ERR_BAD_ARG, "null context");
  return WL_OK;
}

static inline int nchunk_of(const wl_ctx* c) { return (int)((c->g.stride[0] + NT - 1) / NT); }

// x0 sites marched per thread: the largest of 4, 2, 1 dividing L0 that still
// leaves at least 512 workgroups per spinor column.
static int march_for(const wl_ctx* c, int prec, int nrhs) {
  int m = c->march[nrhs > 1 ? 1 : 0][prec];
  if (m > 0) return m;
  const int L0 = c->geom.L[0];
  for (int cand : {4, 2}) {
    if (L0 % cand == 0 && (long long)nchunk_of(c) * (L0 / cand) >= 512) return cand;
  }
  return 1;
}

// x0 sites marched per thread by a multi-RHS launch: only the FP16 two-RHS
// march kernel marches, and only when a march length is set explicitly
static int march_mrhs(const wl_ctx* c, int prec, int k) {
  const int m = c->march[1][prec];
  return (prec == F16 && k == 2 && m > 1) ? m : 1;
}

// right-hand sides per thread: the largest of (4, 2) dividing the number of
// right-hand sides per configuration (FP64: 2), unless set explicitly.
static int mrhs_for(const wl_ctx* c, int prec, int nrhs) {
  if (c->geom.nc != 3) return 1;
  int k = c->mrhs[prec];
  if (k == 0) {
    for (int cand : {4, 2})
      if ((prec != F64 || cand == 2) && nrhs % cand == 0) return cand;
    return 1;
  }
  return (k > 1 && nrhs % k == 0) ? k : 1;
}

// tiles of a one-site-per-thread launch: flat, or L0 * (L1 / slab) * (L2 /
// slab2) slabs of ceil(slab * slab2 * L3 / NT) tiles
static inline int site_tiles(const wl_ctx* c, int slab, int slab2) {
  if (slab <= 0) return (int)((c->V + NT - 1) / NT);
  const long long sites = (long long)slab * slab2 * c->geom.L[3];
  return c->geom.L[0] * (c->geom.L[1] / slab) * (c->geom.L[2] / slab2) * (int)((sites + NT - 1) / NT);
}

// upper bound of reduction partials per column over every kernel (a slab
// order adds at most one partial tile per slab, i.e. at most L0 * L1 * L2)
static inline size_t max_tiles(const wl_ctx* c) {
  size_t a = (size_t)nchunk_of(c) * c->geom.L[0];
  size_t b = (size_t)((c->V + NT - 1) / NT) + (size_t)c->geom.L[0] * c->geom.L[1] * c->geom.L[2];
  // the plane kernel: at most one tile per (x0, x1, x2)
  const size_t d = (size_t)c->geom.L[0] * c->geom.L[1] * c->geom.L[2];
  if (b > a) a = b;
  return a > d ? a : d;
}

// x2 rows of a plane-kernel tile: the largest divisor of L2 whose rows fit one
// workgroup, or 0 when the lattice does not fit the kernel
static int plane_rows(const wl_ctx* c) {
  const int L2 = c->geom.L[2], L3 = c->geom.L[3];
  if (2 * L3 > NT) return 0;
  for (int p = NT / L3; p >= 1; --p)
    if (L2 % p == 0) return p;
  return 0;
}

extern "C" {

int wl_version(void) { return 1; }

const char* wl_last_error(void) { return g_err.c_str(); }

int wl_ctx_create(const wl_geom* geom, const wl_params* params, wl_ctx** out) {
  if (!geom || !params || !out) return fail(WL_ERR_BAD_ARG, "null argument");
  if (geom->nd != 4) return fail(WL_ERR_UNSUPPORTED, "only Nd=4 lattices are supported (got %d)", geom->nd);
  if (geom->nc != 3 && geom->nc != 1) return fail(WL_ERR_UNSUPPORTED, "Nc must be 3 or 1 (got %d)", geom->nc);
  if (geom->batch < 1) return fail(WL_ERR_BAD_ARG, "batch must be >= 1");
  long long V = 1;
  for (int i = 0; i < 4; ++i) {
    if (geom->L[i] < 1) return fail(WL_ERR_BAD_ARG, "lattice extent L[%d]=%d must be >= 1", i, geom->L[i]);
    V *= geom->L[i];
  }
  if (V > (1LL << 31)) return fail(WL_ERR_UNSUPPORTED, "volume %lld too large", V);
  if (std::fabs(params->wilson_r - 1.0) > 1e-12)
    return fail(WL_ERR_UNSUPPORTED, "wilson_r=%g: the half-spinor kernels need r=1", params->wilson_r);
  wl_ctx* c = new wl_ctx();
  c->geom = *geom;
  c->params = *params;
  c->V = V;
  c->Vp = V + PLANE_PAD;
  {
    bool all_even = true;
    for (int i = 0; i < 4; ++i) all_even = all_even && (geom->L[i] % 2 == 0);
    c->Vh = all_even ? V / 2 : 0;
    c->Vhp = all_even ? V / 2 + PLANE_PAD : 0;
  }
  for (int i = 0; i < 4; ++i) c->g.L[i] = geom->L[i];
  c->g.stride[3] = 1;
  c->g.stride[2] = geom->L[3];
  c->g.stride[1] = geom->L[3] * geom->L[2];
  c->g.stride[0] = geom->L[3] * geom->L[2] * geom->L[1];
  c->g.V = V;
  c->g.Vp = c->Vp;
  if (hipGetDevice(&c->device) != hipSuccess ||
      hipDeviceGetAttribute(&c->lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device) !=
          hipSuccess ||
      hipDeviceGetAttribute(&c->lds_per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess) {
    delete c;
    return fail(WL_ERR_HIP, "no HIP device");
  }
  if (hipEventCreateWithFlags(&c->order_ev, hipEventDisableTiming) != hipSuccess) {
    delete c;
    return fail(WL_ERR_HIP, "stream-order event");
  }
  // the device address of wl_apply_noalloc's scratch array, taken once per device
  noalloc_init_device(c->device);
  *out = c;
  return WL_OK;
}

static void prof_release(wl_ctx* c) {
  for (auto& row : c->prof.ev)
    for (auto& e : row)
      if (e) {
        (void)hipEventDestroy(e);
        e = nullptr;
      }
  c->prof.on = false;
}

void wl_ctx_destroy(wl_ctx* c) {
  if (!c) return;
  if (c->order_set) (void)hipEventSynchronize(c->order_ev);
  c->partials.release();
  c->cols.release();
  for (auto& b : c->ws) b.release();
  if (c->host_cols) (void)hipHostFree(c->host_cols);
  if (c->order_ev) (void)hipEventDestroy(c->order_ev);
  prof_release(c);
  delete c;
}

int wl_ctx_set_profiling(wl_ctx* c, int on) {
  CHK(check_ctx(c));
  std::lock_guard<std::mutex> lk(c->mu);
  auto& p = c->prof;
  if (on && !p.on) {
    for (auto& row : p.ev)
      for (auto& e : row)
        if (hipEventCreate(&e) != hipSuccess) {
          e = nullptr;
          prof_release(c);
          return fail(WL_ERR_HIP, "profiling events");
        }
    p.on = true;
  } else if (!on && p.on) {
    prof_release(c);
  }
  p.n = 0;
  p.sum_op[0] = p.sum_op[1] = p.sum_dag[0] = p.sum_dag[1] = p.sum_xpay = 0.0;
  p.cnt[0] = p.cnt[1] = 0;
  p.no_update = 0;
  return WL_OK;
}

int wl_ctx_get_profile(wl_ctx* c, wl_profile* out) {
  CHK(check_ctx(c));
  if (!out) return fail(WL_ERR_BAD_ARG, "null profile");
  std::lock_guard<std::mutex> lk(c->mu);
  const auto& p = c->prof;
  memset(out, 0, sizeof(*out));
  for (int q = 0; q < 2; ++q) {
    out->iterations[q] = p.cnt[q];
    out->us_op[q] = p.cnt[q] ? p.sum_op[q] / p.cnt[q] : 0.0;
    out->us_op_dag[q] = p.cnt[q] ? p.sum_dag[q] / p.cnt[q] : 0.0;
  }
  const int n = p.cnt[0] + p.cnt[1];
  out->us_xpay = n ? p.sum_xpay / n : 0.0;
  out->iterations_without_update = p.no_update;
  return WL_OK;
}

int wl_ctx_set_mass(wl_ctx* c, double mass) {
  CHK(check_ctx(c));
  std::lock_guard<std::mutex> lk(c->mu);
  c->params.mass = mass;
  return WL_OK;
}

// 4D workgroup boxes (b0, b1, b2, b3), product NT = 256, as log2 extents
static const int kBoxShapes[9][4] = {
    {0, 0, 0, 0}, {3, 2, 0, 3}, {2, 2, 1, 3}, {1, 2, 2, 3}, {2, 1, 2, 3},
    {3, 1, 1, 3}, {4, 1, 0, 3}, {2, 2, 0, 4}, {3, 1, 0, 4}};

static bool box_fits(const wl_ctx* c, int shape) {
  if (shape <= 0 || shape > 8) return false;
  for (int mu = 0; mu < 4; ++mu)
    if (c->geom.L[mu] % (1 << kBoxShapes[shape][mu]) != 0) return false;
  return true;
}

// LDS box kernel presets: log2 extents (x0, x1, x2, x3) of the 256-site box
static const int kLdsBoxShapes[4][4] = {{0, 0, 0, 0}, {1, 1, 2, 4}, {1, 2, 2, 3}, {2, 1, 2, 3}};

static bool lds_box_fits(const wl_ctx* c, int shape) {
  if (shape <= 0 || shape > 3 || c->geom.nc != 3) return false;
  for (int mu = 0; mu < 4; ++mu)
    if (c->geom.L[mu] % (1 << kLdsBoxShapes[shape][mu]) != 0) return false;
  return true;
}

// Splits the box grid (bn boxes per direction) into 8 regions, one per XCD:
// halving x0, x1 and x2 when the box counts allow (measured best), else any
// split of 8 over the directions; none (lregions = 0) when the number of
// boxes is not a multiple of 8.
static void lds_box_regions(DslashArgs& a) {
  static const int kSplits[20][4] = {
      {2, 2, 2, 1}, {8, 1, 1, 1}, {2, 2, 1, 2}, {2, 1, 2, 2}, {1, 2, 2, 2}, {4, 2, 1, 1}, {2, 4, 1, 1},
      {4, 1, 2, 1}, {1, 4, 2, 1}, {2, 1, 4, 1}, {1, 2, 4, 1}, {1, 8, 1, 1}, {1, 1, 8, 1}, {4, 1, 1, 2},
      {1, 4, 1, 2}, {1, 1, 4, 2}, {2, 1, 1, 4}, {1, 2, 1, 4}, {1, 1, 2, 4}, {1, 1, 1, 8}};
  a.lregions = 0;
  for (int mu = 0; mu < 4; ++mu) a.lsplit[mu] = 1;
  for (const auto& e : kSplits) {
    bool ok = true;
    for (int mu = 0; mu < 4; ++mu) ok = ok && a.bn[mu] % e[mu] == 0;
    if (!ok) continue;
    for (int mu = 0; mu < 4; ++mu) a.lsplit[mu] = e[mu];
    a.lregions = 1;
    return;
  }
}

int wl_ctx_set_tuning(wl_ctx* c, int prec, int multi_rhs, const wl_tuning* t) {
  CHK(check_ctx(c));
  if (!t) return fail(WL_ERR_BAD_ARG, "null tuning");
  if (prec < F64 || prec > F16) return fail(WL_ERR_BAD_ARG, "bad precision %d", prec);
  if (t->march < 0 || (t->march > 0 && c->geom.L[0] % t->march != 0))
    return fail(WL_ERR_BAD_ARG, "march %d must be 0 (automatic) or divide L0=%d", t->march, c->geom.L[0]);
  if (t->rhs_per_thread != 0 && t->rhs_per_thread != 1 && t->rhs_per_thread != 2 && t->rhs_per_thread != 4)
    return fail(WL_ERR_BAD_ARG, "rhs_per_thread must be 0 (automatic), 1, 2 or 4");
  if (t->block_shape < 0 || t->block_shape > 8 || (t->block_shape > 0 && !box_fits(c, t->block_shape)))
    return fail(WL_ERR_BAD_ARG, "block_shape %d: must be 0 or a preset 1..8 whose extents divide the lattice",
                t->block_shape);
  if (t->blocks_per_cu < 0 || t->blocks_per_cu > 32)
    return fail(WL_ERR_BAD_ARG, "blocks_per_cu must be 0 (no cap) .. 32");
  if (t->cols_outer < 0 || t->cols_outer > 64) return fail(WL_ERR_BAD_ARG, "cols_outer must be 0 .. 64");
  if (t->plane < 0 || (t->plane > 0 && c->geom.L[1] % t->plane != 0))
    return fail(WL_ERR_BAD_ARG, "plane %d must be 0 (off) or divide L1=%d", t->plane, c->geom.L[1]);
  if (t->eo_slab < 0 || (t->eo_slab > 0 && c->geom.L[1] % t->eo_slab != 0))
    return fail(WL_ERR_BAD_ARG, "eo_slab %d must be 0 (flat) or divide L1=%d", t->eo_slab, c->geom.L[1]);
  if (t->slab < 0 || (t->slab > 0 && c->geom.L[1] % t->slab != 0))
    return fail(WL_ERR_BAD_ARG, "slab %d must be 0 (flat) or divide L1=%d", t->slab, c->geom.L[1]);
  if (t->slab_x2 < 0 || (t->slab_x2 > 0 && c->geom.L[2] % t->slab_x2 != 0))
    return fail(WL_ERR_BAD_ARG, "slab_x2 %d must be 0 (whole x1 planes) or divide L2=%d", t->slab_x2, c->geom.L[2]);
  if (t->lds_box < 0 || t->lds_box > 3 || (t->lds_box > 0 && !lds_box_fits(c, t->lds_box)))
    return fail(WL_ERR_BAD_ARG, "lds_box %d: must be 0 or a preset 1..3 whose extents divide the lattice (SU(3) only)",
                t->lds_box);
  std::lock_guard<std::mutex> lk(c->mu);
  const int m = multi_rhs ? 1 : 0;
  c->march[m][prec] = t->march;
  c->xcd_remap[m][prec] = t->xcd_remap ? 1 : 0;
  c->slab[m][prec] = t->slab;
  c->slab2[m][prec] = t->slab_x2;
  c->cols_outer[m][prec] = t->cols_outer;
  c->blocks_per_cu[m][prec] = t->blocks_per_cu;
  c->block_shape[m][prec] = t->block_shape;
  c->plane[m][prec] = t->plane;
  c->eo_slab[m][prec] = t->eo_slab;
  c->lds_box[m][prec] = t->lds_box;
  if (multi_rhs) c->mrhs[prec] = t->rhs_per_thread;
  return WL_OK;
}

int wl_ctx_set_eo_slab(wl_ctx* c, int prec, int multi_rhs, int eo_slab) {
  CHK(check_ctx(c));
  if (prec < F64 || prec > F16) return fail(WL_ERR_BAD_ARG, "bad precision %d", prec);
  if (eo_slab < 0 || (eo_slab > 0 && c->geom.L[1] % eo_slab != 0))
    return fail(WL_ERR_BAD_ARG, "eo_slab %d must be 0 (flat) or divide L1=%d", eo_slab, c->geom.L[1]);
  std::lock_guard<std::mutex> lk(c->mu);
  c->eo_slab[multi_rhs ? 1 : 0][prec] = eo_slab;
  return WL_OK;
}

int wl_ctx_get_tuning(wl_ctx* c, int prec, int nrhs, wl_tuning* t) {
  CHK(check_ctx(c));
  if (!t) return fail(WL_ERR_BAD_ARG, "null tuning");
  if (prec < F64 || prec > F16 || nrhs < 1) return fail(WL_ERR_BAD_ARG, "bad precision or nrhs");
  const int k = mrhs_for(c, prec, nrhs);
  const int m = nrhs > 1 ? 1 : 0;
  t->march = k > 1 ? march_mrhs(c, prec, k) : march_for(c, prec, nrhs);
  t->rhs_per_thread = k;
  t->xcd_remap = c->xcd_remap[m][prec];
  t->slab = t->march == 1 ? c->slab[m][prec] : 0;
  t->slab_x2 = t->slab ? c->slab2[m][prec] : 0;
  t->cols_outer = c->cols_outer[m][prec];
  t->blocks_per_cu = c->blocks_per_cu[m][prec];
  t->block_shape = t->march == 1 ? c->block_shape[m][prec] : 0;
  t->plane = c->plane[m][prec];
  t->eo_slab = c->eo_slab[m][prec];
  t->lds_box = prec == F16 ? c->lds_box[m][prec] : 0;
  return WL_OK;
}

int wl_gauge_upload(wl_ctx* c, const void* U, int layout, int dtype, int where, wl_gauge** out) {
  CHK(check_ctx(c));
  if (!U || !out) return fail(WL_ERR_BAD_ARG, "null argument");
  if (layout != WL_LAYOUT_BMXYIJ && layout != WL_LAYOUT_BXYMIJ) return fail(WL_ERR_BAD_ARG, "bad layout %d", layout);
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  std::lock_guard<std::mutex> lk(c->mu);
  const int nc = c->geom.nc;
  const int ncfg = c->geom.batch;
  size_t src_bytes = (size_t)ncfg * 4 * c->V * nc * nc * (dtype == WL_C128 ? 16 : 8);
  wl_gauge* g = new wl_gauge();
  g->ctx = c;
  g->ncfg = ncfg;
  size_t cb = cfg_bytes(c, F64, 18);
  int rc = g->master.ensure(cb * ncfg);
  if (rc) {
    delete g;
    return rc;
  }
  const void* src = U;
  DevBuf staging;
  if (where == 0) {
    // the upload takes no stream: the links may still be in flight on any
    // stream of the caller (a non-blocking stream does not order against the
    // null stream the pack runs on), so all device work finishes first
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      g->master.release();
      delete g;
      return fail(WL_ERR_HIP, "gauge upload: %s", hipGetErrorString(e));
    }
  }
  if (where == 1) {
    rc = staging.ensure(src_bytes);
    if (rc) {
      g->master.release();
      delete g;
      return rc;
    }
    hipError_t e = hipMemcpy(staging.p, U, src_bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      staging.release();
      g->master.release();
      delete g;
      return fail(WL_ERR_HIP, "gauge upload copy: %s", hipGetErrorString(e));
    }
    src = staging.p;
  }
  hipError_t e = pack_gauge_f64(nc, src, dtype, layout, g->master.p, c->V, c->Vp, ncfg, cb, cb / 4, 0);
  DevBuf dev;
  if (e == hipSuccess && nc == 3) {
    rc = dev.ensure(sizeof(unsigned long long));
    if (rc) {
      staging.release();
      g->master.release();
      delete g;
      return rc;
    }
    e = hipMemsetAsync(dev.p, 0, sizeof(unsigned long long), 0);
    if (e == hipSuccess)
      e = row2_deviation(g->master.p, c->V, c->Vp, ncfg, cb, cb / 4, (unsigned long long*)dev.p, 0);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess && nc == 3) {
    unsigned long long bits = 0;
    e = hipMemcpy(&bits, dev.p, sizeof(bits), hipMemcpyDeviceToHost);
    double d;
    memcpy(&d, &bits, sizeof(d));
    g->row2_dev = (d == d) ? d : 1.0;  // NaN links: never exact
  }
  dev.release();
  staging.release();
  if (e != hipSuccess) {
    g->master.release();
    delete g;
    return fail(WL_ERR_HIP, "gauge pack: %s", hipGetErrorString(e));
  }
  *out = g;
  return WL_OK;
}

void wl_gauge_release(wl_gauge* g) {
  if (!g) return;
  g->master.release();
  for (auto& kv : g->enc) kv.second.release();
  delete g;
}

}  // extern "C"

// A new link encoding is built on the caller's stream but later read from
// whatever stream uses the gauge next, so it is complete before it is handed
// out.  Under graph capture the build would only be recorded, not run:
// refused there.
static int encoding_build_check(hipStream_t st, int prec, int recon) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
    return fail(WL_ERR_BAD_ARG,
                "link encoding (precision %d, recon %d) is built on first use and cannot be built under graph "
                "capture: call wl_gauge_prepare before capturing", prec, recon);
  (void)hipGetLastError();
  return WL_OK;
}

static int encoding_finish(hipStream_t st, hipError_t e, DevBuf& b, const char* what) {
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    b.release();
    return fail(WL_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  }
  return WL_OK;
}

// returns device pointer of the (prec, recon) encoding, building it if needed
static int gauge_encoding(wl_gauge* g, int prec, int recon, hipStream_t st, const void** out) {
  wl_ctx* c = g->ctx;
  if (c->geom.nc == 1) recon = 18;
  if (recon != 18 && recon != 12) return fail(WL_ERR_BAD_ARG, "recon must be 18 or 12 (got %d)", recon);
  if (prec == F64 && recon == 18) {
    *out = g->master.p;
    return WL_OK;
  }
  int key = prec * 100 + recon;
  auto it = g->enc.find(key);
  if (it != g->enc.end()) {
    *out = it->second.p;
    return WL_OK;
  }
  CHK(encoding_build_check(st, prec, recon));
  DevBuf b;
  size_t cb = cfg_bytes(c, prec, recon);
  CHK(b.ensure(cb * g->ncfg));
  size_t mcb = cfg_bytes(c, F64, 18);
  hipError_t e = convert_gauge(prec, c->geom.nc, recon, g->master.p, b.p, c->V, c->Vp, g->ncfg, mcb, mcb / 4, cb, cb / 4, st);
  CHK(encoding_finish(st, e, b, "gauge convert"));
  g->enc[key] = b;
  *out = b.p;
  return WL_OK;
}

// Links per parity in compact order (the even-odd kernels), built from the
// FP64 master on first use.
static int gauge_encoding_eo(wl_gauge* g, int prec, int recon, hipStream_t st, const void** out) {
  wl_ctx* c = g->ctx;
  if (c->Vh == 0) return fail(WL_ERR_UNSUPPORTED, "even-odd operators need every lattice extent even");
  if (c->geom.nc == 1) recon = 18;
  if (recon != 18 && recon != 12) return fail(WL_ERR_BAD_ARG, "recon must be 18 or 12 (got %d)", recon);
  const int key = 1000 + prec * 100 + recon;
  auto it = g->enc.find(key);
  if (it != g->enc.end()) {
    *out = it->second.p;
    return WL_OK;
  }
  CHK(encoding_build_check(st, prec, recon));
  DevBuf b;
  const size_t cb = cfg_bytes_eo(c, prec, recon);
  CHK(b.ensure(cb * g->ncfg));
  const size_t mcb = cfg_bytes(c, F64, 18);
  hipError_t e = convert_gauge_eo(prec, c->geom.nc, recon, g->master.p, b.p, c->g, c->Vh, c->Vhp, g->ncfg, mcb, mcb / 4,
                                  cb, cb / 8, st);
  CHK(encoding_finish(st, e, b, "gauge convert (even-odd)"));
  g->enc[key] = b;
  *out = b.p;
  return WL_OK;
}

// launch geometry once ncol and cols_per_cfg are set
static void set_launch_shape(const wl_ctx* c, DslashArgs& a, int prec, int recon) {
  a.mrhs = mrhs_for(c, prec, a.cols_per_cfg);
  a.march = a.mrhs > 1 ? march_mrhs(c, prec, a.mrhs) : march_for(c, prec, a.cols_per_cfg);
  a.xcd_remap = c->xcd_remap[a.cols_per_cfg > 1 ? 1 : 0][prec];
  a.slab = a.march == 1 ? c->slab[a.cols_per_cfg > 1 ? 1 : 0][prec] : 0;
  {
    const int s2 = c->slab2[a.cols_per_cfg > 1 ? 1 : 0][prec];
    a.slab2 = s2 > 0 ? s2 : c->geom.L[2];
  }
  {
    const int bs = a.march == 1 ? c->block_shape[a.cols_per_cfg > 1 ? 1 : 0][prec] : 0;
    a.box = box_fits(c, bs) ? 1 : 0;
    for (int mu = 0; mu < 4; ++mu) {
      a.bsh[mu] = a.box ? kBoxShapes[bs][mu] : 0;
      a.bn[mu] = c->geom.L[mu] >> a.bsh[mu];
    }
    if (a.box) a.slab = 0;
  }
  {
    // column blocks: the largest block size not above the setting that
    // divides the number of column units (columns, or groups of mrhs)
    const int units = a.ncol / (a.mrhs > 1 ? a.mrhs : 1);
    int cb = c->cols_outer[a.cols_per_cfg > 1 ? 1 : 0][prec];
    if (cb > units) cb = units;
    while (cb > 1 && units % cb != 0) --cb;
    a.cols_outer = cb;
  }
  {
    // LDS reservation that leaves room for at most n resident workgroups per CU
    const int n = c->blocks_per_cu[a.cols_per_cfg > 1 ? 1 : 0][prec];
    long long b = 0;
    if (n > 0 && c->lds_per_cu > 0) {
      b = (long long)c->lds_per_cu / n - 1024;  // the kernels' own reduction scratch is well under 1 KiB
      if (b > c->lds_per_block) b = c->lds_per_block;
      if (b < 0) b = 0;
    }
    a.lds_bytes = (int)b;
  }
  a.slab_tiles = a.slab ? (int)(((long long)a.slab * a.slab2 * c->geom.L[3] + NT - 1) / NT) : 0;
  a.nseg = c->geom.L[0] / a.march;
  a.ntiles = a.march == 1 ? (a.box ? (int)(c->V / NT) : site_tiles(c, a.slab, a.slab2)) : nchunk_of(c) * a.nseg;
  // LDS-staged plane kernel (FP16, one column per thread): replaces the shape above
  a.plane = 0;
  a.prow = 0;
  const int pm = c->plane[a.cols_per_cfg > 1 ? 1 : 0][prec];
  const int prow = plane_rows(c);
  if (pm > 0 && prec == F16 && a.mrhs == 1 && prow > 0 && c->geom.L[1] % pm == 0) {
    a.plane = 1;
    a.prow = prow;
    a.march = pm;
    a.slab = 0;
    a.box = 0;
    a.nseg = c->geom.L[1] / pm;
    a.ntiles = c->geom.L[0] * (c->geom.L[2] / prow) * a.nseg;
    const int need = (prow + 2) * c->geom.L[3] * (32 * c->geom.nc + 8);
    if (a.lds_bytes < need) a.lds_bytes = need;
  }
  // LDS box kernel (F16 SU(3), 12-real links): replaces every shape above
  a.lbox = 0;
  const int lb = c->lds_box[a.cols_per_cfg > 1 ? 1 : 0][prec];
  if (prec == F16 && recon == 12 && lds_box_fits(c, lb)) {
    a.lbox = 1;
    a.plane = 0;
    a.prow = 0;
    a.box = 0;
    a.slab = 0;
    a.slab_tiles = 0;
    a.march = 1;
    a.nseg = c->geom.L[0];
    a.lds_bytes = 0;
    for (int mu = 0; mu < 4; ++mu) {
      a.bsh[mu] = kLdsBoxShapes[lb][mu];
      a.bn[mu] = c->geom.L[mu] >> a.bsh[mu];
    }
    a.ntiles = (int)(c->V / NT);
    lds_box_regions(a);
  }
}

static void fill_dslash_common(const wl_ctx* c, DslashArgs& a, int prec) {
  memset(&a, 0, sizeof(a));
  a.g = c->g;
  for (int i = 0; i < 4; ++i) {
    a.bc_re[i] = c->params.bc_re[i];
    a.bc_im[i] = c->params.bc_im[i];
  }
  a.m0 = c->params.mass + 4.0 * c->params.wilson_r;
  a.r = c->params.wilson_r;
  a.march = 1;  // set_launch_shape fills the launch shape once ncol is known
  a.nseg = c->geom.L[0];
  a.ntiles = (int)((c->V + NT - 1) / NT);
}

// Stencil arguments of the even-odd kernels: compact fields of Vh sites with
// plane stride Vhp, links per parity.
static void fill_eo_args(const wl_ctx* c, DslashArgs& a, int prec, const void* U, int recon, int ncol, int nrhs) {
  fill_dslash_common(c, a, prec);
  const int rc = (c->geom.nc == 1) ? 18 : recon;
  a.U = U;
  a.cfg_bytes = cfg_bytes_eo(c, prec, rc);
  a.dir_bytes = a.cfg_bytes / 8;
  a.ncol = ncol;
  a.cols_per_cfg = nrhs;
  a.g.Vp = c->Vhp;
  a.Vh = c->Vh;
  a.march = 1;
  a.nseg = c->geom.L[0];
  a.box = 0;
  a.cols_outer = 0;
  a.plane = 0;
  a.slab = c->eo_slab[nrhs > 1 ? 1 : 0][prec];
  if (a.slab > 0) {
    const long long plane = (long long)c->geom.L[2] * c->geom.L[3] / 2;
    a.slab_tiles = (int)((a.slab * plane + NT - 1) / NT);
    a.ntiles = c->geom.L[0] * (c->geom.L[1] / a.slab) * a.slab_tiles;
  } else {
    a.slab_tiles = 0;
    a.ntiles = (int)((c->Vh + NT - 1) / NT);
  }
  // two right-hand sides per thread share each link load whenever the
  // right-hand sides per configuration pair up (SU(3))
  a.mrhs = (nrhs > 1 && c->geom.nc == 3 && nrhs % 2 == 0) ? 2 : 1;
  a.xcd_remap = c->xcd_remap[nrhs > 1 ? 1 : 0][prec];
  a.normalized = 1;
}

// op on internal fields of precision prec (no gating, no reduction)
static int dslash_internal(wl_ctx* c, wl_gauge* g, int op, int prec, int recon, const void* in, void* out, int ncol,
                           int nrhs, size_t col_bytes, hipStream_t st) {
  const void* U;
  CHK(gauge_encoding(g, prec, recon, st, &U));
  DslashArgs a;
  fill_dslash_common(c, a, prec);
  int rc = (c->geom.nc == 1) ? 18 : recon;
  a.U = U;
  a.cfg_bytes = cfg_bytes(c, prec, rc);
  a.dir_bytes = a.cfg_bytes / 4;
  a.col_bytes = col_bytes;
  a.ncol = ncol;
  a.cols_per_cfg = nrhs;
  set_launch_shape(c, a, prec, rc);
  a.in = in;
  a.out = out;
  hipError_t e = launch_dslash(prec, c->geom.nc, rc, op, 0, a, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "dslash launch: %s", hipGetErrorString(e));
  return WL_OK;
}

static int check_op(int op) {
  if (op < WL_OP_D || op > WL_OP_DDAGD) return fail(WL_ERR_BAD_ARG, "bad op %d", op);
  return WL_OK;
}

static int op_internal(wl_ctx* c, wl_gauge* g, int op, int prec, int recon, const void* in, void* out, void* tmp,
                       int ncol, int nrhs, size_t col_bytes, hipStream_t st) {
  if (op == WL_OP_DDAGD) {
    CHK(dslash_internal(c, g, WL_OP_D, prec, recon, in, tmp, ncol, nrhs, col_bytes, st));
    return dslash_internal(c, g, WL_OP_DDAG, prec, recon, tmp, out, ncol, nrhs, col_bytes, st);
  }
  return dslash_internal(c, g, op, prec, recon, in, out, ncol, nrhs, col_bytes, st);
}

static int check_apply_args(const wl_ctx* c, const wl_gauge* g, int op, const void* in, const void* out, int nrhs,
                            int dtype) {
  CHK(check_ctx(c));
  if (!g || g->ctx != c) return fail(WL_ERR_BAD_ARG, "gauge handle does not belong to this context");
  if (!in || !out) return fail(WL_ERR_BAD_ARG, "null field pointer");
  if (nrhs < 1) return fail(WL_ERR_BAD_ARG, "nrhs must be >= 1");
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  return check_op(op);
}

static int apply_checked_run(wl_ctx* c, wl_gauge* g, int op, const void* in, void* out, int nrhs, int dtype, void* b0,
                             void* b1, void* b2, hipStream_t st) {
  const int nc = c->geom.nc;
  const int ncol = c->geom.batch * nrhs;
  const size_t cb = col_bytes_of(c, F64);
  hipError_t e = pack_spinor(F64, nc, in, dtype, b0, c->V, c->Vp, ncol, cb, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "pack: %s", hipGetErrorString(e));
  CHK(op_internal(c, g, op, F64, 18, b0, b1, b2, ncol, nrhs, cb, st));
  e = unpack_spinor(F64, nc, b1, out, dtype, c->V, c->Vp, ncol, cb, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "unpack: %s", hipGetErrorString(e));
  return WL_OK;
}

namespace wl {

// Bytes of one of the three FP64 buffers of wl_apply for nrhs right-hand
// sides per configuration.
size_t apply_buffer_bytes(const wl_ctx* c, int nrhs) { return col_bytes_of(c, F64) * (size_t)c->geom.batch * nrhs; }

int ctx_device(const wl_ctx* c) { return c->device; }

// wl_apply: the reference-layout input is packed into b0, op is applied into
// b1 (b2 holds D psi of D^dag D) and b1 is unpacked into out; each buffer
// holds apply_buffer_bytes(c, nrhs).  With b0 == nullptr the buffers are the
// context workspace (allocated on first use, ordered against the context's
// other calls); otherwise they are the caller's and nothing is allocated.
int apply_checked(wl_ctx* c, wl_gauge* g, int op, const void* in, void* out, int nrhs, int dtype, void* b0, void* b1,
                  void* b2, hipStream_t st) {
  CHK(check_apply_args(c, g, op, in, out, nrhs, dtype));
  std::lock_guard<std::mutex> lk(c->mu);
  if (!b0) {
    StreamOrder order(c, st);
    const size_t need = apply_buffer_bytes(c, nrhs);
    CHK(c->ws[0].ensure(need));
    CHK(c->ws[1].ensure(need));
    CHK(c->ws[2].ensure(need));
    return apply_checked_run(c, g, op, in, out, nrhs, dtype, c->ws[0].p, c->ws[1].p, c->ws[2].p, st);
  }
  return apply_checked_run(c, g, op, in, out, nrhs, dtype, b0, b1, b2, st);
}

}  // namespace wl

extern "C" {

int wl_apply(wl_ctx* c, wl_gauge* g, int op, const void* in, void* out, int nrhs, int dtype, void* stream) {
  return wl::apply_checked(c, g, op, in, out, nrhs, dtype, nullptr, nullptr, nullptr, (hipStream_t)stream);
}

int wl_gauge_prepare(wl_gauge* g, int prec, int recon, void* stream) {
  if (!g) return fail(WL_ERR_BAD_ARG, "null gauge");
  if (prec < F64 || prec > F16) return fail(WL_ERR_BAD_ARG, "bad precision %d", prec);
  std::lock_guard<std::mutex> lk(g->ctx->mu);
  const void* U;
  return gauge_encoding(g, prec, recon, (hipStream_t)stream, &U);
}

int wl_field_create(wl_ctx* c, int prec, int ncol, wl_field** out) {
  CHK(check_ctx(c));
  if (!out || ncol < 1) return fail(WL_ERR_BAD_ARG, "bad field arguments");
  if (prec < F64 || prec > F16) return fail(WL_ERR_BAD_ARG, "bad precision %d", prec);
  wl_field* f = new wl_field();
  f->ctx = c;
  f->prec = prec;
  f->ncol = ncol;
  f->col_bytes = col_bytes_of(c, prec);
  int rc = f->buf.ensure(f->col_bytes * ncol);
  if (rc) {
    delete f;
    return rc;
  }
  hipError_t e = hipMemset(f->buf.p, 0, f->col_bytes * ncol);
  if (e != hipSuccess) {
    f->buf.release();
    delete f;
    return fail(WL_ERR_HIP, "field clear: %s", hipGetErrorString(e));
  }
  *out = f;
  return WL_OK;
}

void wl_field_destroy(wl_field* f) {
  if (!f) return;
  f->buf.release();
  delete f;
}

size_t wl_field_bytes(const wl_field* f) { return f ? f->col_bytes * f->ncol : 0; }

int wl_field_import(wl_field* f, const void* src, int dtype, void* stream) {
  if (!f || !src) return fail(WL_ERR_BAD_ARG, "null argument");
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  wl_ctx* c = f->ctx;
  // ordered with the context's other calls like every call on it (a field
  // written by a stencil on one stream and read here on another)
  std::lock_guard<std::mutex> lk(c->mu);
  StreamOrder order(c, (hipStream_t)stream);
  hipError_t e = pack_spinor(f->prec, c->geom.nc, src, dtype, f->buf.p, c->V, c->Vp, f->ncol, f->col_bytes, (hipStream_t)stream);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "import: %s", hipGetErrorString(e));
  return WL_OK;
}

int wl_field_export(const wl_field* f, void* dst, int dtype, void* stream) {
  if (!f || !dst) return fail(WL_ERR_BAD_ARG, "null argument");
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  wl_ctx* c = f->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  StreamOrder order(c, (hipStream_t)stream);
  hipError_t e = unpack_spinor(f->prec, c->geom.nc, f->buf.p, dst, dtype, c->V, c->Vp, f->ncol, f->col_bytes, (hipStream_t)stream);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "export: %s", hipGetErrorString(e));
  return WL_OK;
}

int wl_dslash(wl_ctx* c, wl_gauge* g, int op, int recon, const wl_field* in, wl_field* out, int nrhs, void* stream) {
  CHK(check_ctx(c));
  if (!g || g->ctx != c || !in || !out || in->ctx != c || out->ctx != c)
    return fail(WL_ERR_BAD_ARG, "handles do not belong to this context");
  if (in->prec != out->prec || in->ncol != out->ncol) return fail(WL_ERR_BAD_ARG, "in/out fields differ in precision or columns");
  if (in == out) return fail(WL_ERR_BAD_ARG, "in and out must differ");
  if (nrhs < 1 || in->ncol != c->geom.batch * nrhs) return fail(WL_ERR_BAD_ARG, "field has %d columns, expected batch*nrhs", in->ncol);
  CHK(check_op(op));
  std::lock_guard<std::mutex> lk(c->mu);
  StreamOrder order(c, (hipStream_t)stream);
  void* tmp = nullptr;
  if (op == WL_OP_DDAGD) {
    CHK(c->ws[9].ensure(in->col_bytes * in->ncol));
    tmp = c->ws[9].p;
  }
  return op_internal(c, g, op, in->prec, recon, in->buf.p, out->buf.p, tmp, in->ncol, nrhs, in->col_bytes,
                     (hipStream_t)stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// CG on D^dag D with reliable updates
// ---------------------------------------------------------------------------
namespace {

struct Solver {
  wl_ctx* c;
  wl_gauge* g;
  hipStream_t st;
  int nc, ncol, nrhs, ntiles;
  int P, R;       // inner precision, recon
  int PY;         // precision of the inner solution accumulator
  bool mixed;
  bool fused;     // reliable updates run inside every iteration (device-driven)
  bool rec = false;  // mixed precision: search direction by recurrence (WL_SEARCH_RECURRENCE): pb[0]
                     // holds the column's p, t holds D p, and both are updated in place
  bool eo = false;  // operator M = the normalized even Schur complement S_e (else M = D)
  size_t cb64, cbP, cbY;
  const void* U64;
  int R64 = 18;    // link encoding of U64 (12 when it reproduces the stored links exactly)
  const void* UP;
  void *b, *x;    // F64 caller-side internal fields
  void *rhi, *thi;
  void *r, *t, *y;
  void* pb[2];    // search directions, used in turn: iteration k reads pb[k % 2] and writes pb[(k + 1) % 2]
  void *mid64 = nullptr, *midP = nullptr;  // odd-parity intermediates of S_e (even-odd only)
  Reduction red{};  // copied into every launch's arguments: fields not set stay zero
  wl_ctx::Prof* prof = nullptr;  // kernel timing of host-driven iterations (null: off)

  // the four events of the next timed iteration, or null
  hipEvent_t* prof_slot(int par) {
    if (!prof || !prof->on || prof->n >= wl_ctx::Prof::SLOTS) return nullptr;
    prof->par[prof->n] = par;
    return prof->ev[prof->n++];
  }

  void dargs(DslashArgs& a, int prec, const void* U, int recon) {
    if (eo) {
      fill_eo_args(c, a, prec, U, recon, ncol, nrhs);
    } else {
      fill_dslash_common(c, a, prec);
      a.U = U;
      a.cfg_bytes = cfg_bytes(c, prec, nc == 1 ? 18 : recon);
      a.dir_bytes = a.cfg_bytes / 4;
      a.ncol = ncol;
      a.cols_per_cfg = nrhs;
      set_launch_shape(c, a, prec, nc == 1 ? 18 : recon);
    }
    a.m0_from_state = 1;
    a.red = red;
  }
  BlasArgs bargs(int gate) {
    BlasArgs a;
    memset(&a, 0, sizeof(a));
    a.V = eo ? c->Vh : c->V;
    a.Vp = eo ? c->Vhp : c->Vp;
    a.ncol = ncol;
    a.ntiles = ntiles;
    a.gate = gate;
    a.red = red;
    return a;
  }

  // out = M in (dag = false) or M^dag in, with epilogue epi and reduction
  // action; d carries the remaining arguments (gate, w, r, y).  The even-odd
  // operator is two launches through the odd intermediate mid.
  int apply(DslashArgs d, int prec, int recon, bool dag, const void* in, void* out, void* mid, int epi, int action) {
    hipError_t e;
    if (!eo) {
      d.in = in;
      d.out = out;
      d.red.action = action;
      e = launch_dslash(prec, nc, nc == 1 ? 18 : recon, dag ? WL_OP_DDAG : WL_OP_D, epi, d, st);
      if (e != hipSuccess) return fail(WL_ERR_HIP, "dslash: %s", hipGetErrorString(e));
      return WL_OK;
    }
    DslashArgs h = d;
    h.pt = 1;
    h.in = in;
    h.out = mid;
    h.red.action = RA_NONE;
    e = launch_eo(prec, nc, recon, dag ? 1 : 0, 0, EPI_STORE, h, st);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "even-odd hop: %s", hipGetErrorString(e));
    d.pt = 0;
    d.in = mid;
    d.ctr = in;
    d.out = out;
    d.red.action = action;
    e = launch_eo(prec, nc, recon, dag ? 1 : 0, 1, epi, d, st);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "even-odd Schur: %s", hipGetErrorString(e));
    return WL_OK;
  }

  // x += y for the columns the gate admits, including an alpha p still held
  // back by a pairwise column
  hipError_t accumulate(BlasArgs a) {
    a.a = x; a.col_bytes_a = cb64;
    a.b = y; a.col_bytes_b = cbY;
    a.c = pb[0]; a.d = rec ? pb[0] : pb[1]; a.col_bytes_c = cbP; a.col_bytes_d = cbP;
    return launch_accumulate(F64, PY, P, nc, a, st);
  }

  // gated reliable update before an iteration that reads pb[par ^ 1]:
  // x += y; r = b - A x (FP64 arithmetic, stored in the inner precision);
  // pb[par ^ 1] = r + beta p (p: each column's latest search direction).
  // With the search direction by recurrence the next iteration's launches
  // form p = r + beta p and D p from the new r themselves.
  int reliable_update(int par) {
    hipError_t e;
    if (mixed) {
      e = accumulate(bargs(2));
      if (e != hipSuccess) return fail(WL_ERR_HIP, "accumulate: %s", hipGetErrorString(e));
    }
    DslashArgs d;
    dargs(d, F64, U64, R64);
    d.gate = 2;
    d.col_bytes = cb64;
    CHK(apply(d, F64, R64, false, x, thi, mid64, EPI_STORE, RA_NONE));
    d.w = b;
    // with the search direction by recurrence nothing runs between the true
    // residual and the update's bookkeeping, so its reduction closes the
    // update (one launch fewer)
    d.red.then_finalize = rec ? 1 : 0;
    if (mixed) {
      // r = b - A x straight into the inner precision (no FP64 copy of it)
      d.res = r;
      d.col_bytes_y = cbP;
      CHK(apply(d, F64, R64, true, thi, nullptr, mid64, P == F32 ? EPI_RESID32 : EPI_RESID16, RA_TRUERES));
    } else {
      CHK(apply(d, F64, R64, true, thi, rhi, mid64, EPI_RESID, RA_TRUERES));
    }
    if (rec) return WL_OK;
    CHK(xpay(bargs(2), par, true));
    e = launch_finalize_update(red.cols, ncol, st);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "finalize: %s", hipGetErrorString(e));
    return WL_OK;
  }

  // pb[par ^ 1] = r + beta p for the columns the gate admits, p = pb[par]
  // (an iteration) or each column's latest search direction (an update,
  // whose column may have stopped iterating in an earlier iteration)
  int xpay(BlasArgs a, int par, bool latest) {
    a.a = latest ? pb[0] : pb[par]; a.col_bytes_a = cbP;
    a.d = latest ? pb[1] : nullptr;
    a.b = r; a.col_bytes_b = cbP;
    a.c = pb[par ^ 1];
    hipError_t e = launch_xpay(P, nc, a, st);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "xpay: %s", hipGetErrorString(e));
    return WL_OK;
  }

  // The first update of a solve from x = 0: r_hi = b - A 0 = b exactly, so
  // the FP64 operator applications are skipped: r = b (in the inner
  // precision), |r_hi|^2 = |b|^2 decides convergence, p = r (into pb[0],
  // where the first iteration reads it; pb[1] is zero and beta is 0; with the
  // search direction by recurrence beta = 0 makes the first iteration's
  // launches set p = r and t = D r).
  int initial_update_zero_guess() {
    hipError_t e;
    {
      BlasArgs a = bargs(2);
      a.a = r; a.col_bytes_a = mixed ? cbP : cb64;
      a.b = b; a.col_bytes_b = cb64;
      e = launch_convert(mixed ? P : F64, F64, nc, a, st);
      if (e != hipSuccess) return fail(WL_ERR_HIP, "convert: %s", hipGetErrorString(e));
    }
    e = launch_trueres_from_b2(red.cols, ncol, rec ? 1 : 0, st);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "true residual: %s", hipGetErrorString(e));
    if (rec) return WL_OK;
    CHK(xpay(bargs(2), 1, true));
    e = launch_finalize_update(red.cols, ncol, st);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "finalize: %s", hipGetErrorString(e));
    return WL_OK;
  }

  // One CG iteration on A = D^dag D, two stencil launches with fused vector
  // work: t = D p with |t|^2 = <p, A p> (alpha), then A p = D^dag t consumed
  // in the epilogue by r -= alpha A p, |r|^2 (beta) and the y update; then
  // p' = r + beta p into the other buffer.  The iteration of parity par reads
  // p = pb[par]; the previous iteration's p, in pb[par ^ 1], lets a pairwise
  // column add alpha p of two iterations to y at once (every second
  // iteration), which halves the y traffic of the epilogue.
  //
  // With the search direction by recurrence (rec) the iteration is two
  // launches: t = D r + beta t_prev (= D p for p = r + beta p_prev, t_prev
  // read at the site itself) with |t|^2, then D^dag t with p = r + beta p_prev
  // formed and stored in place beside r -= alpha A p, |r|^2 and the y update.
  // No pass of its own for p, and p is read only at the site itself.
  int iteration(int par, bool with_update = true) {
    hipEvent_t* ev = prof_slot(par);
    DslashArgs d;
    dargs(d, P, UP, R);
    d.gate = 1;
    d.col_bytes = cbP;
    d.red.par = par;
    d.search_rec = rec ? 1 : 0;
    if (ev) HIPCHK(hipEventRecord(ev[0], st));
    CHK(apply(d, P, R, false, rec ? r : pb[par], t, midP, EPI_NORM, RA_PAP));
    if (ev) HIPCHK(hipEventRecord(ev[1], st));
    d.w = pb[par]; d.w2 = pb[par ^ 1]; d.res = r; d.y = y; d.col_bytes_y = cbY;
    const int epi = PY == F64 ? EPI_CG64 : (PY == F32 ? EPI_CG32 : EPI_CG16);
    CHK(apply(d, P, R, true, t, rec ? pb[0] : nullptr, midP, epi, RA_ITER));
    if (ev) HIPCHK(hipEventRecord(ev[2], st));
    if (!rec) CHK(xpay(bargs(1), par, false));
    if (ev) HIPCHK(hipEventRecord(ev[3], st));
    // Mixed precision: the reliable update runs on the device right after the
    // iteration that requested it (all its kernels are gated on the column's
    // need_update flag), so no host round trip is needed.
    if (fused && with_update) return reliable_update(par);
    return WL_OK;
  }
};

}  // namespace

// Reliable-update trigger actually used: the requested delta, but not below
// a floor set by the inner storage precision (an iterated residual kept in
// FP16 with per-site scales cannot follow a drop much below a few 1e-3 of its
// starting value, and the trigger would then never fire).
static double effective_delta(double delta, int inner) {
  if (!(delta > 0.0)) return 0.0;
  const double floor = inner == F16 ? 3e-3 : (inner == F32 ? 1e-5 : 0.0);
  return delta < floor ? floor : delta;
}

// storage of the inner solution accumulator y (x itself for a plain FP64 solve)
static int accum_prec_of(const wl_solve_params* sp, int inner) {
  if (inner == F64) return F64;
  switch (sp->accum_prec) {
    case WL_ACCUM_F64: return F64;
    case WL_ACCUM_F16: return F16;
    default: return F32;
  }
}

// Adds the timed iterations recorded since the last call (the stream has
// been synchronized) to the context's kernel-time sums.
static void prof_harvest(wl_ctx::Prof& p) {
  for (int k = 0; k < p.n; ++k) {
    float t01 = 0.f, t12 = 0.f, t23 = 0.f;
    if (hipEventElapsedTime(&t01, p.ev[k][0], p.ev[k][1]) != hipSuccess ||
        hipEventElapsedTime(&t12, p.ev[k][1], p.ev[k][2]) != hipSuccess ||
        hipEventElapsedTime(&t23, p.ev[k][2], p.ev[k][3]) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    const int q = p.par[k] & 1;
    p.sum_op[q] += 1e3 * t01;
    p.sum_dag[q] += 1e3 * t12;
    p.sum_xpay += 1e3 * t23;
    p.cnt[q] += 1;
  }
  p.n = 0;
}

// CG on M^dag M x = b for internal FP64 fields b, x: M = D (full lattice) or,
// with eo, M = the normalized even Schur complement on compact even fields.
static int cg_fields(wl_ctx* c, wl_gauge* g, void* b, void* x, int use_x0, int nrhs, const wl_solve_params* sp,
                     wl_solve_stats* stats, hipStream_t st, bool eo = false) {
  const int nc = c->geom.nc;
  const int ncol = c->geom.batch * nrhs;
  if (sp->inner_prec < F64 || sp->inner_prec > F16) return fail(WL_ERR_BAD_ARG, "bad inner precision %d", sp->inner_prec);
  if (sp->maxiter < 0) return fail(WL_ERR_BAD_ARG, "maxiter must be >= 0");
  if (!(sp->tol >= 0.0)) return fail(WL_ERR_BAD_ARG, "tol must be >= 0");
  if (sp->use_graph && st == nullptr) return fail(WL_ERR_BAD_ARG, "graph capture needs a non-default stream");
  Solver s;
  s.c = c;
  s.g = g;
  s.st = st;
  s.nc = nc;
  s.ncol = ncol;
  s.nrhs = nrhs;
  s.eo = eo;
  s.ntiles = (int)(((eo ? c->Vh : c->V) + NT - 1) / NT);
  if (sp->search_update != WL_SEARCH_RECURRENCE && sp->search_update != WL_SEARCH_XPAY)
    return fail(WL_ERR_BAD_ARG, "bad search_update %d", sp->search_update);
  s.P = sp->inner_prec;
  s.R = (nc == 1) ? 18 : (sp->inner_recon == 12 ? 12 : 18);
  s.mixed = s.P != F64;
  s.fused = s.mixed;
  s.rec = s.mixed && sp->search_update == WL_SEARCH_RECURRENCE;
  s.PY = accum_prec_of(sp, s.P);
  s.cb64 = eo ? col_bytes_eo(c, F64) : col_bytes_of(c, F64);
  s.cbP = eo ? col_bytes_eo(c, s.P) : col_bytes_of(c, s.P);
  s.cbY = eo ? col_bytes_eo(c, s.PY) : col_bytes_of(c, s.PY);
  s.R64 = update_recon(g);
  if (eo) {
    CHK(gauge_encoding_eo(g, F64, s.R64, st, &s.U64));
    CHK(gauge_encoding_eo(g, s.P, s.R, st, &s.UP));
  } else {
    CHK(gauge_encoding(g, F64, s.R64, st, &s.U64));
    CHK(gauge_encoding(g, s.P, s.R, st, &s.UP));
  }
  s.b = b;
  s.x = x;
  // workspace
  if (!s.mixed) CHK(c->ws[0].ensure(s.cb64 * ncol));  // r_hi (a mixed solve stores r straight in the inner precision)
  CHK(c->ws[1].ensure(s.cb64 * ncol));  // t_hi
  CHK(c->ws[2].ensure(s.cbP * ncol));   // p (buffer 0)
  CHK(c->ws[11].ensure(s.cbP * ncol));  // p (buffer 1)
  CHK(c->ws[4].ensure(s.cbP * ncol));   // t
  s.rhi = s.mixed ? nullptr : c->ws[0].p;
  s.thi = c->ws[1].p;
  s.pb[0] = c->ws[2].p;
  s.pb[1] = c->ws[11].p;
  s.t = c->ws[4].p;
  if (eo) {
    CHK(c->ws[3].ensure(s.cbP * ncol));    // odd intermediate, inner precision
    CHK(c->ws[10].ensure(s.cb64 * ncol));  // odd intermediate, FP64
    s.midP = c->ws[3].p;
    s.mid64 = c->ws[10].p;
  }
  if (s.mixed) {
    CHK(c->ws[5].ensure(s.cbP * ncol));  // r
    CHK(c->ws[6].ensure(s.cbY * ncol));  // y
    s.r = c->ws[5].p;
    s.y = c->ws[6].p;
  } else {
    s.r = s.rhi;
    s.y = s.x;
  }
  CHK(c->partials.ensure(sizeof(double) * (size_t)ncol * max_tiles(c)));
  CHK(c->cols.ensure(sizeof(ColState) * ncol));
  if (c->host_cols_n < (size_t)ncol) {
    if (c->host_cols) (void)hipHostFree(c->host_cols);
    c->host_cols = nullptr;
    c->host_cols_n = 0;
    HIPCHK(hipHostMalloc((void**)&c->host_cols, sizeof(ColState) * ncol, hipHostMallocDefault));
    c->host_cols_n = ncol;
  }
  s.red.partials = (double*)c->partials.p;
  s.red.cols = (ColState*)c->cols.p;
  s.red.action = RA_NONE;
  // kernel timing runs the iterations host-driven, with events between launches
  s.prof = c->prof.on ? &c->prof : nullptr;
  const bool use_graph = sp->use_graph && !s.prof;

  HIPCHK(hipMemsetAsync(s.pb[0], 0, s.cbP * ncol, st));
  HIPCHK(hipMemsetAsync(s.pb[1], 0, s.cbP * ncol, st));
  if (s.mixed) HIPCHK(hipMemsetAsync(s.y, 0, s.cbY * ncol, st));
  if (!use_x0) HIPCHK(hipMemsetAsync(x, 0, s.cb64 * ncol, st));

  ColState* hc = c->host_cols;
  for (int i = 0; i < ncol; ++i) {
    ColState& cs = hc[i];
    memset(&cs, 0, sizeof(cs));
    double m = sp->mass ? sp->mass[i] : c->params.mass;
    cs.m0 = m + 4.0 * c->params.wilson_r;
    cs.maxiter = sp->maxiter;
    cs.mode = sp->stop_rule == WL_STOP_ITERATED ? 1 : 0;
    cs.delta = s.mixed ? effective_delta(sp->delta, s.P) : 0.0;
    cs.pairwise = s.mixed ? 1 : 0;
    cs.p_par = 1;  // pb[1] is zero: the first update sets p = r in pb[0]
    cs.need_update = 1;
    cs.active = 0;
    cs.status = WL_STATUS_RUNNING;
  }
  HIPCHK(hipMemcpyAsync(c->cols.p, hc, sizeof(ColState) * ncol, hipMemcpyHostToDevice, st));
  {
    BlasArgs a = s.bargs(0);
    a.red.action = RA_B2;
    a.a = b;
    a.col_bytes_a = s.cb64;
    hipError_t e = launch_norm2(F64, nc, a, st);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "norm: %s", hipGetErrorString(e));
  }
  HIPCHK(hipMemcpyAsync(hc, c->cols.p, sizeof(ColState) * ncol, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  const double tol2 = sp->tol * sp->tol;
  for (int i = 0; i < ncol; ++i) {
    hc[i].tol2b2 = tol2 * hc[i].b2;
    hc[i].stop2 = tol2 * hc[i].b2;
  }
  HIPCHK(hipMemcpyAsync(c->cols.p, hc, sizeof(ColState) * ncol, hipMemcpyHostToDevice, st));
  CHK(use_x0 ? s.reliable_update(1) : s.initial_update_zero_guess());

  // Iteration blocks of `poll`, 4 and 2 iterations (as HIP graphs when
  // requested); after each block the host reads the per-column state and
  // picks the next block from the predicted number of remaining iterations
  // (the observed residual decay rate), so little work is queued past
  // convergence.  Every block has an even length, so each one starts on
  // search-direction buffer 0 and ends having written it.
  const int poll = sp->poll_every > 0 ? (sp->poll_every + 1) / 2 * 2 : 8;
  std::vector<int> sizes;
  for (int n : {poll, 4, 2})
    if (n <= poll && (sizes.empty() || n < sizes.back())) sizes.push_back(n);
  std::vector<hipGraph_t> graphs(sizes.size(), nullptr);
  std::vector<hipGraphExec_t> execs(sizes.size(), nullptr);
  auto cleanup = [&]() {
    for (auto& e : execs)
      if (e) (void)hipGraphExecDestroy(e);
    for (auto& g : graphs)
      if (g) (void)hipGraphDestroy(g);
  };
  if (use_graph) {
    for (size_t k = 0; k < sizes.size(); ++k) {
      hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) {
        cleanup();
        return fail(WL_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
      }
      int rc = WL_OK;
      for (int i = 0; i < sizes[k] && rc == WL_OK; ++i) rc = s.iteration(i & 1);
      e = hipStreamEndCapture(st, &graphs[k]);
      if (rc != WL_OK) {
        cleanup();
        return rc;
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&execs[k], graphs[k], nullptr, nullptr, 0);
      if (e != hipSuccess) {
        cleanup();
        return fail(WL_ERR_HIP, "graph capture/instantiate: %s", hipGetErrorString(e));
      }
    }
  }
  std::vector<double> prev_rr(ncol, 0.0);
  std::vector<int> prev_it(ncol, 0);
  int result = WL_OK;
  long long guard = 0;
  const long long guard_max = 64LL + 8LL * sp->maxiter;
  while (true) {
    // (errors break out of the loop so that the captured graphs are freed)
    hipError_t pe = hipMemcpyAsync(hc, c->cols.p, sizeof(ColState) * ncol, hipMemcpyDeviceToHost, st);
    if (pe == hipSuccess) pe = hipStreamSynchronize(st);
    if (pe != hipSuccess) {
      result = fail(WL_ERR_HIP, "solver state read: %s", hipGetErrorString(pe));
      break;
    }
    if (s.prof) prof_harvest(*s.prof);
    bool all_done = true, any_update = false, any_active = false;
    double remaining = 0.0;
    bool have_rate = false;
    // Mixed precision: the iterations of the next block at which some active
    // column may ask for a reliable update -- its residual drop reaching
    // delta, the forced update after MAX_INNER_ITERS, its stopping threshold
    // or maxiter -- predicted from its observed decay with margins (a 25%
    // faster decay than observed, two iterations early, three late): each
    // column's first request h and, delta later, its next one h + T.  A column
    // without a rate yet may ask at any iteration.
    std::vector<char> may_update(sizes[0], 0);
    auto mark = [&](double from, double to) {
      const int a = std::max(0, (int)std::floor(from) - 2);
      const int b = std::min((int)may_update.size() - 1, (int)std::ceil(to) + 3);
      for (int j = a; j <= b; ++j) may_update[j] = 1;
    };
    double horizon = 1e30;
    for (int i = 0; i < ncol; ++i) {
      const ColState& cs = hc[i];
      if (!cs.done) all_done = false;
      if (cs.need_update && !cs.done) any_update = true;
      if (!cs.active) continue;
      any_active = true;
      // the inner stopping threshold as the device applies it (with the
      // margin of the last true / iterated residual ratio)
      const double stop2 = cs.mode == 1 ? cs.stop2 : cs.stop2 / std::max(1.0, cs.gap2);
      if (cs.iter > prev_it[i] && prev_rr[i] > cs.rr && cs.rr > 0.0 && stop2 > 0.0) {
        double rate = std::log(prev_rr[i] / cs.rr) / (cs.iter - prev_it[i]);
        double left = cs.rr > stop2 ? std::log(cs.rr / stop2) / rate : 1.0;
        left = std::min(left, (double)(cs.maxiter - cs.iter));
        remaining = std::max(remaining, left);
        have_rate = true;
        const double fast = 1.25 * rate;
        double h = cs.rr > stop2 ? std::log(cs.rr / stop2) / fast : 0.0;
        if (cs.delta > 0.0) {
          const double tgt = cs.delta * cs.delta * cs.r0norm * cs.r0norm;
          h = std::min(h, cs.rr > tgt ? std::log(cs.rr / tgt) / fast : 0.0);
        }
        h = std::min(h, (double)(MAX_INNER_ITERS - (cs.iter - cs.last_update)));
        h = std::min(h, (double)(cs.maxiter - cs.iter));
        horizon = std::min(horizon, h);
        mark(h, h);
        // after an update the residual restarts near its iterated value: the
        // next request is a drop of delta (or the stopping threshold) later
        double T = cs.delta > 0.0 ? std::log(1.0 / (cs.delta * cs.delta)) / fast : 1e30;
        const double to_stop = cs.rr > stop2 ? std::log(cs.rr / stop2) / fast - h : 0.0;
        T = std::min(T, std::max(0.0, to_stop));
        T = std::min(T, (double)MAX_INNER_ITERS);
        // the later requests of the block: windows every T (T >= 1), and
        // every iteration from the first one where T cannot be trusted
        if (T >= 4.0) {
          for (double t = h + T; t < (double)may_update.size() + 3.0; t += T) mark(t, t);
        } else {
          mark(h + T, 1e9);
        }
      } else {
        remaining = std::max(remaining, (double)poll);
        horizon = 0.0;
        mark(0.0, 1e9);
      }
      prev_rr[i] = cs.rr;
      prev_it[i] = cs.iter;
    }
    if (all_done) break;
    if (++guard > guard_max) {
      result = fail(WL_ERR_HIP, "solver made no progress");
      break;
    }
    if (any_update) {
      result = s.reliable_update(1);
      if (result != WL_OK) break;
      continue;
    }
    if (!any_active) {
      result = fail(WL_ERR_HIP, "solver state inconsistent: no active column");
      break;
    }
    size_t k = 0;
    if (have_rate) {
      while (k + 1 < sizes.size() && sizes[k] > remaining + 1.0) ++k;
    }
    // iterations at which no column can ask for an update run without the
    // fused (gated, then idle) update launches; a column that asks outside the
    // predicted windows keeps its state (it stops iterating) until the next
    // fused update of the block or the host's update after the block, so the
    // iterates, and the solution, are the same as with every update fused
    int skipped = 0;
    if (s.fused)
      for (int i = 0; i < sizes[k]; ++i) skipped += may_update[i] ? 0 : 1;
    if (skipped > 0) {
      if (s.prof) s.prof->no_update += skipped;
      for (int i = 0; i < sizes[k] && result == WL_OK; ++i) result = s.iteration(i & 1, may_update[i] != 0);
      if (result != WL_OK) break;
    } else if (execs[k]) {
      hipError_t e = hipGraphLaunch(execs[k], st);
      if (e != hipSuccess) {
        result = fail(WL_ERR_HIP, "graph launch: %s", hipGetErrorString(e));
        break;
      }
    } else {
      for (int i = 0; i < sizes[k] && result == WL_OK; ++i) result = s.iteration(i & 1);
      if (result != WL_OK) break;
    }
  }
  cleanup();
  if (result != WL_OK) return result;
  if (s.mixed) {
    // fold any remaining inner correction (columns stopped by NaN/breakdown)
    hipError_t e = s.accumulate(s.bargs(0));
    if (e != hipSuccess) return fail(WL_ERR_HIP, "accumulate: %s", hipGetErrorString(e));
  }
  if (stats) {
    for (int i = 0; i < ncol; ++i) {
      const ColState& cs = hc[i];
      double nb = cs.b2 > 0 ? std::sqrt(cs.b2) : 1.0;
      stats[i].iters = cs.iter;
      stats[i].status = cs.status;
      stats[i].reliable_updates = cs.nupdates;
      stats[i].pad = 0;
      stats[i].true_resid = cs.b2 > 0 ? std::sqrt(cs.rtrue2) / nb : std::sqrt(cs.rtrue2);
      stats[i].iter_resid = cs.b2 > 0 ? std::sqrt(cs.rr) / nb : std::sqrt(cs.rr);
    }
  }
  return WL_OK;
}

extern "C" {

int wl_cg_normal_field(wl_ctx* c, wl_gauge* g, const wl_field* b, wl_field* x, int use_x0, int nrhs,
                       const wl_solve_params* sp, wl_solve_stats* stats, void* stream) {
  CHK(check_ctx(c));
  if (!g || g->ctx != c || !b || !x || !sp) return fail(WL_ERR_BAD_ARG, "null or foreign handle");
  if (b->prec != F64 || x->prec != F64) return fail(WL_ERR_BAD_ARG, "b and x must be FP64 fields");
  if (nrhs < 1 || b->ncol != c->geom.batch * nrhs || x->ncol != b->ncol)
    return fail(WL_ERR_BAD_ARG, "fields must have batch*nrhs columns");
  std::lock_guard<std::mutex> lk(c->mu);
  StreamOrder order(c, (hipStream_t)stream);
  return cg_fields(c, g, b->buf.p, x->buf.p, use_x0, nrhs, sp, stats, (hipStream_t)stream);
}

int wl_cg_normal(wl_ctx* c, wl_gauge* g, const void* b, void* x, const void* x0, int nrhs, int dtype,
                 const wl_solve_params* sp, wl_solve_stats* stats, void* stream) {
  CHK(check_ctx(c));
  if (!g || g->ctx != c || !b || !x || !sp) return fail(WL_ERR_BAD_ARG, "null or foreign handle");
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  if (nrhs < 1) return fail(WL_ERR_BAD_ARG, "nrhs must be >= 1");
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  StreamOrder order(c, st);
  const int nc = c->geom.nc;
  const int ncol = c->geom.batch * nrhs;
  const size_t cb = col_bytes_of(c, F64);
  CHK(c->ws[7].ensure(cb * ncol));
  CHK(c->ws[8].ensure(cb * ncol));
  hipError_t e = pack_spinor(F64, nc, b, dtype, c->ws[7].p, c->V, c->Vp, ncol, cb, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "pack: %s", hipGetErrorString(e));
  if (x0) {
    e = pack_spinor(F64, nc, x0, dtype, c->ws[8].p, c->V, c->Vp, ncol, cb, st);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "pack: %s", hipGetErrorString(e));
  }
  CHK(cg_fields(c, g, c->ws[7].p, c->ws[8].p, x0 != nullptr, nrhs, sp, stats, st));
  e = unpack_spinor(F64, nc, c->ws[8].p, x, dtype, c->V, c->Vp, ncol, cb, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "unpack: %s", hipGetErrorString(e));
  return WL_OK;
}

int wl_eo_apply(wl_ctx* c, wl_gauge* g, int op, int normalized, const void* in, void* out, int nrhs, int dtype,
                void* stream) {
  CHK(check_ctx(c));
  if (!g || g->ctx != c) return fail(WL_ERR_BAD_ARG, "gauge handle does not belong to this context");
  if (!in || !out) return fail(WL_ERR_BAD_ARG, "null field pointer");
  if (nrhs < 1) return fail(WL_ERR_BAD_ARG, "nrhs must be >= 1");
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  if (op < WL_EO_K_OE || op > WL_EO_SCHUR_O_DAG) return fail(WL_ERR_BAD_ARG, "bad even-odd op %d", op);
  if (c->Vh == 0) return fail(WL_ERR_UNSUPPORTED, "even-odd operators need every lattice extent even");
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  StreamOrder order(c, st);
  const int nc = c->geom.nc;
  const int ncol = c->geom.batch * nrhs;
  const size_t cb = col_bytes_eo(c, F64);
  const void* U;
  CHK(gauge_encoding_eo(g, F64, 18, st, &U));
  CHK(c->ws[0].ensure(cb * ncol));
  CHK(c->ws[1].ensure(cb * ncol));
  CHK(c->ws[2].ensure(cb * ncol));
  // source parity of the input and target parity of the output
  const bool hop = op <= WL_EO_KDAG_EO;
  const bool dag = (op == WL_EO_KDAG_OE || op == WL_EO_KDAG_EO || op == WL_EO_SCHUR_E_DAG || op == WL_EO_SCHUR_O_DAG);
  int target;
  if (hop) target = (op == WL_EO_K_OE || op == WL_EO_KDAG_OE) ? 1 : 0;
  else target = (op == WL_EO_SCHUR_E || op == WL_EO_SCHUR_E_DAG) ? 0 : 1;
  const int source = hop ? 1 - target : target;
  hipError_t e = pack_spinor_eo(F64, nc, in, dtype, c->ws[0].p, c->g, c->Vh, c->Vhp, source, ncol, cb, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "pack: %s", hipGetErrorString(e));
  DslashArgs d;
  fill_eo_args(c, d, F64, U, 18, ncol, nrhs);
  d.col_bytes = cb;
  d.normalized = normalized ? 1 : 0;
  d.pt = hop ? target : 1 - target;
  d.in = c->ws[0].p;
  d.out = hop ? c->ws[1].p : c->ws[2].p;
  e = launch_eo(F64, nc, 18, dag ? 1 : 0, 0, EPI_STORE, d, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "even-odd hop: %s", hipGetErrorString(e));
  if (!hop) {
    d.pt = target;
    d.in = c->ws[2].p;
    d.ctr = c->ws[0].p;
    d.out = c->ws[1].p;
    e = launch_eo(F64, nc, 18, dag ? 1 : 0, 1, EPI_STORE, d, st);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "even-odd Schur: %s", hipGetErrorString(e));
  }
  e = unpack_spinor_eo(F64, nc, c->ws[1].p, out, dtype, c->g, c->Vhp, target, ncol, cb, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "unpack: %s", hipGetErrorString(e));
  return WL_OK;
}

int wl_pf_force(wl_ctx* c, wl_gauge* g, const void* X, const void* Y, void* F, int layout, int dtype, void* stream) {
  CHK(check_ctx(c));
  if (!g || g->ctx != c) return fail(WL_ERR_BAD_ARG, "gauge handle does not belong to this context");
  if (!X || !Y || !F) return fail(WL_ERR_BAD_ARG, "null field pointer");
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  if (layout != WL_LAYOUT_BMXYIJ && layout != WL_LAYOUT_BXYMIJ) return fail(WL_ERR_BAD_ARG, "bad layout %d", layout);
  std::lock_guard<std::mutex> lk(c->mu);
  ForceArgs a;
  memset(&a, 0, sizeof(a));
  a.g = c->g;
  a.U = g->master.p;
  a.cfg_bytes = cfg_bytes(c, F64, 18);
  a.dir_bytes = a.cfg_bytes / 4;
  a.X = X;
  a.Y = Y;
  a.F = F;
  a.layout = layout;
  for (int i = 0; i < 4; ++i) {
    a.bc_re[i] = c->params.bc_re[i];
    a.bc_im[i] = c->params.bc_im[i];
  }
  a.r = c->params.wilson_r;
  hipError_t e = launch_pf_force(c->geom.nc, dtype, a, c->geom.batch, (hipStream_t)stream);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "force: %s", hipGetErrorString(e));
  return WL_OK;
}

int wl_pf_force_field(wl_ctx* c, wl_gauge* g, const wl_field* X, const wl_field* Y, void* F, int layout, int dtype,
                      void* stream) {
  CHK(check_ctx(c));
  if (!g || g->ctx != c || !X || !Y || X->ctx != c || Y->ctx != c)
    return fail(WL_ERR_BAD_ARG, "null or foreign handle");
  if (!F) return fail(WL_ERR_BAD_ARG, "null force pointer");
  if (X->prec != F64 || Y->prec != F64) return fail(WL_ERR_BAD_ARG, "X and Y must be FP64 fields");
  if (X->ncol != c->geom.batch || Y->ncol != c->geom.batch)
    return fail(WL_ERR_BAD_ARG, "X and Y must have one column per configuration (%d)", c->geom.batch);
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  if (layout != WL_LAYOUT_BMXYIJ && layout != WL_LAYOUT_BXYMIJ) return fail(WL_ERR_BAD_ARG, "bad layout %d", layout);
  std::lock_guard<std::mutex> lk(c->mu);
  StreamOrder order(c, (hipStream_t)stream);  // X, Y may come from a call on another stream
  ForceArgs a;
  memset(&a, 0, sizeof(a));
  a.g = c->g;
  a.U = g->master.p;
  a.cfg_bytes = cfg_bytes(c, F64, 18);
  a.dir_bytes = a.cfg_bytes / 4;
  a.X = X->buf.p;
  a.Y = Y->buf.p;
  a.F = F;
  a.layout = layout;
  a.col_bytes = X->col_bytes;
  for (int i = 0; i < 4; ++i) {
    a.bc_re[i] = c->params.bc_re[i];
    a.bc_im[i] = c->params.bc_im[i];
  }
  a.r = c->params.wilson_r;
  hipError_t e = launch_pf_force(c->geom.nc, dtype, a, c->geom.batch, (hipStream_t)stream);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "force: %s", hipGetErrorString(e));
  return WL_OK;
}

int wl_cg_eo_normal(wl_ctx* c, wl_gauge* g, const void* b, void* x, const void* x0, int nrhs, int dtype,
                    const wl_solve_params* sp, wl_solve_stats* stats, void* stream) {
  CHK(check_ctx(c));
  if (!g || g->ctx != c || !b || !x || !sp) return fail(WL_ERR_BAD_ARG, "null or foreign handle");
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  if (nrhs < 1) return fail(WL_ERR_BAD_ARG, "nrhs must be >= 1");
  if (c->Vh == 0) return fail(WL_ERR_UNSUPPORTED, "even-odd operators need every lattice extent even");
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  StreamOrder order(c, st);
  const int nc = c->geom.nc;
  const int ncol = c->geom.batch * nrhs;
  const size_t cb = col_bytes_eo(c, F64);
  CHK(c->ws[7].ensure(cb * ncol));
  CHK(c->ws[8].ensure(cb * ncol));
  hipError_t e = pack_spinor_eo(F64, nc, b, dtype, c->ws[7].p, c->g, c->Vh, c->Vhp, 0, ncol, cb, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "pack: %s", hipGetErrorString(e));
  if (x0) {
    e = pack_spinor_eo(F64, nc, x0, dtype, c->ws[8].p, c->g, c->Vh, c->Vhp, 0, ncol, cb, st);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "pack: %s", hipGetErrorString(e));
  }
  CHK(cg_fields(c, g, c->ws[7].p, c->ws[8].p, x0 != nullptr, nrhs, sp, stats, st, true));
  e = unpack_spinor_eo(F64, nc, c->ws[8].p, x, dtype, c->g, c->Vhp, 0, ncol, cb, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "unpack: %s", hipGetErrorString(e));
  return WL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Continuous-batching solve engine: a persistent set of `capacity` columns on
// one resident gauge configuration.  Right-hand sides are admitted into free
// columns and retired when converged while the others keep iterating; one
// captured iteration graph serves every occupancy because all kernels gate
// per column on the device state (free columns are "done").
// ---------------------------------------------------------------------------
struct wl_batch {
  wl_ctx* c = nullptr;
  wl_gauge* g = nullptr;
  hipStream_t st = nullptr;
  int cap = 0;
  int poll = 8;
  double tol2 = 0.0;
  wl_solve_params sp{};
  Solver s{};
  DevBuf bufs[11];
  DevBuf partials, cols;
  ColState* host = nullptr;
  std::vector<char> busy;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::mutex mu;
};

static size_t batch_bytes_per_col(const wl_ctx* c, int prec) {
  const bool mixed = prec != F64;
  size_t b = 3 * col_bytes_of(c, F64) + 3 * col_bytes_of(c, prec);  // b x t_hi; two p, t
  if (mixed) b += col_bytes_of(c, prec) + col_bytes_of(c, F64);  // r and an accumulator of at most FP64
  else b += col_bytes_of(c, F64);                                 // r_hi (the FP64 iterated residual)
  return b + sizeof(double) * max_tiles(c) + sizeof(ColState);
}

static void batch_free(wl_batch* bt) {
  if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
  if (bt->graph) (void)hipGraphDestroy(bt->graph);
  for (auto& b : bt->bufs) b.release();
  bt->partials.release();
  bt->cols.release();
  if (bt->host) (void)hipHostFree(bt->host);
  delete bt;
}

extern "C" {

size_t wl_batch_bytes(wl_ctx* c, int capacity, int inner_prec) {
  if (!c || capacity < 1 || inner_prec < F64 || inner_prec > F16) return 0;
  return (size_t)capacity * batch_bytes_per_col(c, inner_prec);
}

int wl_batch_create(wl_ctx* c, wl_gauge* g, int capacity, const wl_solve_params* sp, void* stream, wl_batch** out) {
  CHK(check_ctx(c));
  if (!g || g->ctx != c || !sp || !out) return fail(WL_ERR_BAD_ARG, "null or foreign handle");
  if (c->geom.batch != 1) return fail(WL_ERR_BAD_ARG, "a solve batch needs a context with one gauge configuration");
  if (capacity < 1) return fail(WL_ERR_BAD_ARG, "capacity must be >= 1");
  if (sp->inner_prec < F64 || sp->inner_prec > F16) return fail(WL_ERR_BAD_ARG, "bad inner precision %d", sp->inner_prec);
  if (!(sp->tol >= 0.0) || sp->maxiter < 0) return fail(WL_ERR_BAD_ARG, "bad tol / maxiter");
  if (sp->search_update != WL_SEARCH_RECURRENCE && sp->search_update != WL_SEARCH_XPAY)
    return fail(WL_ERR_BAD_ARG, "bad search_update %d", sp->search_update);
  if (stream == nullptr && sp->use_graph) return fail(WL_ERR_BAD_ARG, "graph capture needs a non-default stream");
  std::lock_guard<std::mutex> lk(c->mu);
  wl_batch* bt = new wl_batch();
  bt->c = c;
  bt->g = g;
  bt->st = (hipStream_t)stream;
  bt->cap = capacity;
  bt->sp = *sp;
  bt->sp.mass = nullptr;
  bt->poll = sp->poll_every > 0 ? (sp->poll_every + 1) / 2 * 2 : 8;  // even: blocks start on search-direction buffer 0
  bt->tol2 = sp->tol * sp->tol;
  bt->busy.assign(capacity, 0);
  Solver& s = bt->s;
  const int nc = c->geom.nc;
  s.c = c;
  s.g = g;
  s.st = bt->st;
  s.nc = nc;
  s.ncol = capacity;
  s.nrhs = capacity;
  s.ntiles = (int)((c->V + NT - 1) / NT);
  s.P = sp->inner_prec;
  s.R = (nc == 1) ? 18 : (sp->inner_recon == 12 ? 12 : 18);
  s.mixed = s.P != F64;
  s.fused = true;
  s.rec = s.mixed && sp->search_update == WL_SEARCH_RECURRENCE;
  s.PY = accum_prec_of(sp, s.P);
  s.cb64 = col_bytes_of(c, F64);
  s.cbP = col_bytes_of(c, s.P);
  s.cbY = col_bytes_of(c, s.PY);
  int rc = WL_OK;
  auto alloc = [&](int i, size_t bytes) -> void* {
    if (rc == WL_OK) rc = bt->bufs[i].ensure(bytes);
    return bt->bufs[i].p;
  };
  s.b = alloc(0, s.cb64 * capacity);
  s.x = alloc(1, s.cb64 * capacity);
  s.rhi = s.mixed ? nullptr : alloc(2, s.cb64 * capacity);
  s.thi = alloc(3, s.cb64 * capacity);
  s.pb[0] = alloc(4, s.cbP * capacity);
  s.pb[1] = alloc(5, s.cbP * capacity);
  s.t = alloc(6, s.cbP * capacity);
  if (s.mixed) {
    s.r = alloc(7, s.cbP * capacity);
    s.y = alloc(8, s.cbY * capacity);
  } else {
    s.r = s.rhi;
    s.y = s.x;
  }
  if (rc == WL_OK) rc = bt->partials.ensure(sizeof(double) * (size_t)capacity * max_tiles(c));
  if (rc == WL_OK) rc = bt->cols.ensure(sizeof(ColState) * capacity);
  s.R64 = update_recon(g);
  if (rc == WL_OK) rc = gauge_encoding(g, F64, s.R64, bt->st, &s.U64);
  if (rc == WL_OK) rc = gauge_encoding(g, s.P, s.R, bt->st, &s.UP);
  if (rc != WL_OK) {
    batch_free(bt);
    return rc;
  }
  if (hipHostMalloc((void**)&bt->host, sizeof(ColState) * capacity, hipHostMallocDefault) != hipSuccess) {
    batch_free(bt);
    return fail(WL_ERR_OOM, "pinned host state");
  }
  s.red.partials = (double*)bt->partials.p;
  s.red.cols = (ColState*)bt->cols.p;
  s.red.action = RA_NONE;
  s.red.only_col = -1;
  for (int i = 0; i < capacity; ++i) {
    memset(&bt->host[i], 0, sizeof(ColState));
    bt->host[i].done = 1;
    bt->host[i].status = WL_STATUS_RUNNING;
  }
  hipError_t e = hipMemcpyAsync(bt->cols.p, bt->host, sizeof(ColState) * capacity, hipMemcpyHostToDevice, bt->st);
  for (int i = 0; i < 9 && e == hipSuccess; ++i)
    if (bt->bufs[i].p) e = hipMemsetAsync(bt->bufs[i].p, 0, bt->bufs[i].n, bt->st);
  if (e == hipSuccess) e = hipStreamSynchronize(bt->st);
  if (e != hipSuccess) {
    batch_free(bt);
    return fail(WL_ERR_HIP, "batch init: %s", hipGetErrorString(e));
  }
  if (sp->use_graph) {
    e = hipStreamBeginCapture(bt->st, hipStreamCaptureModeThreadLocal);
    int r2 = WL_OK;
    for (int i = 0; i < bt->poll && r2 == WL_OK && e == hipSuccess; ++i) r2 = s.iteration(i & 1);
    hipError_t e2 = hipStreamEndCapture(bt->st, &bt->graph);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&bt->exec, bt->graph, nullptr, nullptr, 0);
    if (r2 != WL_OK || e != hipSuccess) {
      batch_free(bt);
      return r2 != WL_OK ? r2 : fail(WL_ERR_HIP, "batch graph: %s", hipGetErrorString(e));
    }
  }
  *out = bt;
  return WL_OK;
}

void wl_batch_destroy(wl_batch* bt) {
  if (!bt) return;
  (void)hipStreamSynchronize(bt->st);
  batch_free(bt);
}

int wl_batch_admit(wl_batch* bt, int slot, const void* b, int dtype, double mass) {
  if (!bt || !b) return fail(WL_ERR_BAD_ARG, "null argument");
  if (slot < 0 || slot >= bt->cap) return fail(WL_ERR_BAD_ARG, "slot %d out of range", slot);
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  std::lock_guard<std::mutex> lk(bt->mu);
  if (bt->busy[slot]) return fail(WL_ERR_BAD_ARG, "slot %d is occupied", slot);
  wl_ctx* c = bt->c;
  Solver& s = bt->s;
  const int nc = c->geom.nc;
  hipStream_t st = bt->st;
  hipError_t e = pack_spinor(F64, nc, b, dtype, (char*)s.b + slot * s.cb64, c->V, c->Vp, 1, s.cb64, st);
  if (e == hipSuccess) e = hipMemsetAsync((char*)s.x + slot * s.cb64, 0, s.cb64, st);
  if (e == hipSuccess) e = hipMemsetAsync((char*)s.pb[0] + slot * s.cbP, 0, s.cbP, st);
  if (e == hipSuccess) e = hipMemsetAsync((char*)s.pb[1] + slot * s.cbP, 0, s.cbP, st);
  if (e == hipSuccess && s.mixed) e = hipMemsetAsync((char*)s.y + slot * s.cbY, 0, s.cbY, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "admit: %s", hipGetErrorString(e));
  ColState& cs = bt->host[slot];
  memset(&cs, 0, sizeof(cs));
  cs.m0 = mass + 4.0 * c->params.wilson_r;
  cs.maxiter = bt->sp.maxiter;
  cs.mode = bt->sp.stop_rule == WL_STOP_ITERATED ? 1 : 0;
  cs.delta = s.mixed ? effective_delta(bt->sp.delta, s.P) : 0.0;
  cs.pairwise = s.mixed ? 1 : 0;
  cs.p_par = 1;
  cs.status = WL_STATUS_RUNNING;
  e = hipMemcpyAsync((ColState*)bt->cols.p + slot, &cs, sizeof(ColState), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "admit: %s", hipGetErrorString(e));
  BlasArgs a = s.bargs(3);
  a.red.only_col = slot;
  a.red.action = RA_B2;
  a.a = s.b;
  a.col_bytes_a = s.cb64;
  e = launch_norm2(F64, nc, a, st);
  if (e == hipSuccess) e = launch_admit_finish((ColState*)bt->cols.p, slot, bt->tol2, st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "admit: %s", hipGetErrorString(e));
  // x = 0: the column's first true residual is b (only this slot awaits an
  // update between iteration blocks, and the update kernels are gated on it)
  CHK(s.initial_update_zero_guess());
  bt->busy[slot] = 1;
  return WL_OK;
}

int wl_batch_step(wl_batch* bt) {
  if (!bt) return fail(WL_ERR_BAD_ARG, "null batch");
  std::lock_guard<std::mutex> lk(bt->mu);
  if (bt->exec) {
    hipError_t e = hipGraphLaunch(bt->exec, bt->st);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "graph launch: %s", hipGetErrorString(e));
    return WL_OK;
  }
  for (int i = 0; i < bt->poll; ++i) CHK(bt->s.iteration(i & 1));
  return WL_OK;
}

int wl_batch_query(wl_batch* bt, wl_solve_stats* stats, int* state) {
  if (!bt) return fail(WL_ERR_BAD_ARG, "null batch");
  std::lock_guard<std::mutex> lk(bt->mu);
  hipError_t e = hipMemcpyAsync(bt->host, bt->cols.p, sizeof(ColState) * bt->cap, hipMemcpyDeviceToHost, bt->st);
  if (e == hipSuccess) e = hipStreamSynchronize(bt->st);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "query: %s", hipGetErrorString(e));
  for (int i = 0; i < bt->cap; ++i) {
    const ColState& cs = bt->host[i];
    if (state) state[i] = !bt->busy[i] ? 0 : (cs.done ? 2 : 1);
    if (stats) {
      double nb = cs.b2 > 0 ? std::sqrt(cs.b2) : 1.0;
      stats[i].iters = cs.iter;
      stats[i].status = cs.status;
      stats[i].reliable_updates = cs.nupdates;
      stats[i].pad = 0;
      stats[i].true_resid = std::sqrt(cs.rtrue2) / nb;
      stats[i].iter_resid = std::sqrt(cs.rr) / nb;
    }
  }
  return WL_OK;
}

int wl_batch_retire(wl_batch* bt, int slot, void* x, int dtype) {
  if (!bt || !x) return fail(WL_ERR_BAD_ARG, "null argument");
  if (slot < 0 || slot >= bt->cap) return fail(WL_ERR_BAD_ARG, "slot %d out of range", slot);
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  std::lock_guard<std::mutex> lk(bt->mu);
  if (!bt->busy[slot]) return fail(WL_ERR_BAD_ARG, "slot %d is free", slot);
  wl_ctx* c = bt->c;
  Solver& s = bt->s;
  // a column stopped early (NaN, breakdown) may hold an unfolded correction
  if (s.mixed) {
    BlasArgs a = s.bargs(3);
    a.red.only_col = slot;
    hipError_t e = s.accumulate(a);
    if (e != hipSuccess) return fail(WL_ERR_HIP, "retire: %s", hipGetErrorString(e));
  }
  hipError_t e = unpack_spinor(F64, c->geom.nc, (char*)s.x + slot * s.cb64, x, dtype, c->V, c->Vp, 1, s.cb64, bt->st);
  if (e == hipSuccess) {
    ColState& cs = bt->host[slot];
    memset(&cs, 0, sizeof(cs));
    cs.done = 1;
    cs.status = WL_STATUS_RUNNING;
    e = hipMemcpyAsync((ColState*)bt->cols.p + slot, &cs, sizeof(ColState), hipMemcpyHostToDevice, bt->st);
  }
  if (e != hipSuccess) return fail(WL_ERR_HIP, "retire: %s", hipGetErrorString(e));
  bt->busy[slot] = 0;
  return WL_OK;
}

int wl_lime_decode(const void* payload, size_t nbytes, int word_bytes, const int* dims, int nd, int datacount,
                   int ncomplex, int order, void* out, int dtype, void* stream) {
  if (!payload || !out || !dims) return fail(WL_ERR_BAD_ARG, "null argument");
  if (word_bytes != 4 && word_bytes != 8) return fail(WL_ERR_BAD_ARG, "word_bytes must be 4 or 8, got %d", word_bytes);
  if (dtype != WL_C128 && dtype != WL_C64) return fail(WL_ERR_BAD_ARG, "bad dtype %d", dtype);
  if (order < WL_LIME_LEXI || order > WL_LIME_COUNT) return fail(WL_ERR_BAD_ARG, "bad order %d", order);
  if (nd < 1 || nd > 4) return fail(WL_ERR_UNSUPPORTED, "lattices of 1 to 4 dimensions, got %d", nd);
  if (datacount < 1 || ncomplex < 1) return fail(WL_ERR_BAD_ARG, "datacount and ncomplex must be >= 1");
  size_t V = 1;
  for (int i = 0; i < nd; ++i) {
    if (dims[i] < 1) return fail(WL_ERR_BAD_ARG, "dims[%d] = %d", i, dims[i]);
    V *= (size_t)dims[i];
  }
  const size_t Wc = (size_t)datacount * (size_t)ncomplex;
  if (Wc > (size_t)1 << 24 || V > (size_t)1 << 36) return fail(WL_ERR_UNSUPPORTED, "field too large");
  const size_t want = V * Wc * 2 * (size_t)word_bytes;
  if (nbytes != want) return fail(WL_ERR_BAD_ARG, "payload holds %zu bytes, the field needs %zu", nbytes, want);
  if (((uintptr_t)payload) % (2 * word_bytes) != 0) return fail(WL_ERR_BAD_ARG, "payload not aligned to one complex value");
  const int osz = dtype == WL_C128 ? 16 : 8;
  if (((uintptr_t)out) % osz != 0) return fail(WL_ERR_BAD_ARG, "output not aligned to one complex value");
  LimeArgs a;
  memset(&a, 0, sizeof(a));
  a.in = payload;
  a.out = out;
  // pad to four axes with x0 first and the last axis in x3
  int L[4] = {1, 1, 1, 1};
  L[0] = dims[0];
  if (nd >= 2) L[3] = dims[nd - 1];
  if (nd >= 3) L[1] = dims[1];
  if (nd == 4) L[2] = dims[2];
  for (int i = 0; i < 4; ++i) a.L[i] = L[i];
  a.V = (long long)V;
  a.Wc = (int)Wc;
  a.item = ncomplex;
  a.dc = datacount;
  a.order = order;
  // tile of up to 16 x 16 site records shrunk to 20 KiB of LDS, x0 first:
  // the output runs along x3 (n3 items of a direction for links) stay long,
  // and several workgroups share a CU (a 32^4 FP64 link field: 4 x 8)
  const size_t rec = Wc * 2 * (size_t)word_bytes;
  int T0 = std::min(L[0], 16), T3 = std::min(L[3], 16);
  while ((size_t)T0 * T3 * rec > 20480 && (T0 > 1 || T3 > 1)) {
    if (T0 > 1 && 2 * T0 >= T3)
      T0 = (T0 + 1) / 2;
    else
      T3 = (T3 + 1) / 2;
  }
  const size_t lds = (size_t)T0 * T3 * rec;
  if (lds > 65536) return fail(WL_ERR_UNSUPPORTED, "a site record of %zu bytes does not fit in LDS", rec);
  a.T0 = T0;
  a.T3 = T3;
  a.nt0 = (L[0] + T0 - 1) / T0;
  a.nt3 = (L[3] + T3 - 1) / T3;
  const long long nblk = (long long)a.nt0 * a.nt3 * L[1] * L[2];
  if (nblk > 0x7fffffffLL) return fail(WL_ERR_UNSUPPORTED, "field too large");
  hipError_t e = launch_lime_decode(word_bytes, dtype == WL_C128, a, lds, (hipStream_t)stream);
  if (e != hipSuccess) return fail(WL_ERR_HIP, "lime decode: %s", hipGetErrorString(e));
  return WL_OK;
}

}  // extern "C"
