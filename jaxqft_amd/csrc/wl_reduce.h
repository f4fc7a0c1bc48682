// Deterministic grid reductions for the CG scalars.
//
// Every workgroup writes one FP64 partial (wave64 shuffle tree + LDS across
// the 4 wavefronts) with a plain store; a one-workgroup-per-column finish
// kernel then sums the partials in a fixed order and performs the per-column
// scalar update.  No atomics or fences (an agent-scope fence per workgroup
// writes back the XCD's L2 and costs more than the stencil itself).  The
// result is independent of workgroup scheduling, so a column gives the same
// bits alone or batched with other columns.
#pragma once
#include "wl_device.h"
#include "wl_kernels.h"

#define WL_STATUS_RUNNING (-1)
#define WL_STATUS_CONVERGED 0
#define WL_STATUS_MAXITER 1
#define WL_STATUS_NAN 2
#define WL_STATUS_BREAKDOWN 3

namespace wl {

__device__ __forceinline__ bool gate_skip(int gate, const ColState& s) {
  if (gate == 1) return !s.active;
  if (gate == 2) return !(s.need_update && !s.done);
  return false;
}

// gate 3 selects one column (admission of a new right-hand side)
__device__ __forceinline__ bool gate_skip_col(int gate, const Reduction& red, int col) {
  if (gate == 3) return col != red.only_col;
  return gate_skip(gate, red.cols[col]);
}

// Whether the gate admits any of the ncol columns: every lane reads the state
// of columns lane, lane + 64, ... and the wavefront votes, so each wavefront
// of a workgroup reaches the same answer with one round of loads.
__device__ __forceinline__ bool gate_any_col(int gate, const Reduction& red, int ncol) {
  bool any = false;
  for (int c = (int)(threadIdx.x & 63u); c < ncol; c += 64) any |= !gate_skip_col(gate, red, c);
  return __ballot(any) != 0ull;
}

__device__ __forceinline__ void col_finish(ColState& s, int status) {
  s.status = status;
  s.done = 1;
  s.active = 0;
}

__device__ __forceinline__ void apply_action(int action, ColState& s, double sum, int par = 0) {
  switch (action) {
    case RA_PAP: {
      s.pAp = sum;
      if (!(sum > 0.0) || !isfinite(sum)) {
        col_finish(s, isfinite(sum) ? WL_STATUS_BREAKDOWN : WL_STATUS_NAN);
        s.alpha = 0.0;
      } else {
        s.alpha = s.rr / sum;
      }
      break;
    }
    case RA_ITER: {
      s.p_par = par;
      if (s.pairwise) {
        // the launch just finished held alpha p back (first of a pair) or added it (second)
        s.alpha_prev = s.alpha;
        s.pend ^= 1;
      }
      s.rr_old = s.rr;
      s.rr = sum;
      s.beta = sum / s.rr_old;
      s.iter += 1;
      // Stopping on the true residual, the iterated one must reach the target
      // with the margin by which it last fell short of the true one (the
      // inner precision's rounding separates them): the convergence check
      // that follows then rarely fails and costs an extra update.
      const double gap2 = s.mode == 1 ? 1.0 : fmax(1.0, s.gap2);
      if (!isfinite(sum)) {
        col_finish(s, WL_STATUS_NAN);
      } else if (sum * gap2 <= s.stop2) {
        if (s.mode == 1) {
          col_finish(s, WL_STATUS_CONVERGED);
        } else {
          s.need_update = 2;
          s.active = 0;
        }
      } else if (s.delta > 0.0 && (sqrt(sum) < s.delta * s.r0norm || s.iter - s.last_update >= MAX_INNER_ITERS)) {
        s.need_update = 1;
        s.active = 0;
      } else if (s.iter >= s.maxiter) {
        // one last true-residual evaluation decides the status
        s.need_update = 2;
        s.active = 0;
      }
      break;
    }
    case RA_B2:
      s.b2 = sum;
      break;
    case RA_TRUERES: {
      s.rtrue2 = sum;
      // rr is still the iterated residual that asked for this update (0 at
      // the start of a solve)
      if (s.rr > 0.0 && isfinite(sum)) s.gap2 = sum / s.rr;
      if (!isfinite(sum)) {
        col_finish(s, WL_STATUS_NAN);
      } else if (sum <= s.tol2b2) {
        col_finish(s, WL_STATUS_CONVERGED);
      } else if (s.iter >= s.maxiter) {
        col_finish(s, WL_STATUS_MAXITER);
      } else {
        // continue from the true residual, keeping the search direction
        s.beta = (s.rr_old > 0.0) ? sum / s.rr_old : 0.0;
      }
      break;
    }
    default:
      break;
  }
}

// Closes a reliable update of a column that asked for one: the iterated
// residual restarts from the true one (a held-back alpha p went into the
// update's x += y), and the column iterates again unless it has finished.
__device__ __forceinline__ void finalize_col(ColState& s) {
  if (!s.need_update) return;
  s.pend = 0;
  if (!s.done) {
    s.rr = s.rtrue2;
    s.r0norm = sqrt(s.rtrue2);
    s.nupdates += 1;
    s.last_update = s.iter;
    s.active = 1;
  }
  s.need_update = 0;
}

// Phase 1, called by every thread of the workgroup: write this workgroup's
// FP64 partial (plain store; the kernel boundary publishes it).
__device__ __forceinline__ void write_partial(const Reduction& red, int col, int blk, int nblk, double v) {
  __shared__ double lds[NT / 64];
  double bs = block_sum<NT>(v, lds);
  if (threadIdx.x == 0) red.partials[(size_t)col * nblk + blk] = bs;
}

// Phase 2 kernel (one workgroup per column): fixed-order sum of the partials
// and the per-column scalar update.  gate: 0 all columns, 1 active columns,
// 2 columns awaiting a reliable update -- the same gate as the phase-1 kernel,
// evaluated on the same (unchanged) state.
__global__ __launch_bounds__(NT) void k_reduce_finish(Reduction red, int nblk, int gate);

}  // namespace wl
