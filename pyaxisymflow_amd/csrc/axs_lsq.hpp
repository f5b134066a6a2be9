// Least-squares extrapolation sweep shared by the drop-in entry point and
// the soft-body stepper.
#pragma once

#include "axs_common.hpp"

namespace axs {

struct LsqWork {
  double* nx;          // fitted eta_x of this sweep, (n0, n1)
  double* ny;          // fitted eta_y of this sweep, (n0, n1)
  int8_t* pend;        // 1 where a cell was fitted this sweep
  unsigned* counters;  // cells fitted per sweep (one slot per sweep; the tail's count after them)
  int* list;           // cells with current ^ target set (list form), (n0, n1) ints
  unsigned* count;     // entries of list
  int16_t* cur2;       // second flag and value buffers of the ping-pong sweeps
  double* ex2;
  double* ey2;
  int* tail;           // the tail's four lists (unknown and just fitted, twice), list capacity each
  int cap;             // list capacity (entries)
};

// One sweep (fit + commit).  Sweep `sweep` exits at once when
// counters[sweep - 1] is zero, so many sweeps can be queued without a sync.
hipError_t launch_lsq_sweep(int16_t* cur, const int16_t* tgt, double* ex, double* ey,
                            const double* gx, const double* gy, int n0, int n1, LsqWork w,
                            int sweep, hipStream_t s);
// List form: launch_lsq_list collects the cells with current ^ target set
// (*w.count must be zero; the caller clears it after the last sweep), then
// each launch_lsq_sweep_list sweeps only those cells, with the same results
// as launch_lsq_sweep on the whole grid.
hipError_t launch_lsq_list(const int16_t* cur, const int16_t* tgt, int64_t n, LsqWork w,
                           hipStream_t s);
hipError_t launch_lsq_sweep_list(int16_t* cur, const int16_t* tgt, double* ex, double* ey,
                                 const double* gx, const double* gy, int n0, int n1, LsqWork w,
                                 int sweep, hipStream_t s);
// One sweep over the listed cells in one launch: reads flags and values from
// (cur_in, ex_in, ey_in), writes (cur_out, ex_out, ey_out) for every listed
// cell (fitted, or copied through); cells off the list must already be equal
// in both sets.  After the loop has ended (the previous sweep fitted
// nothing) a sweep returns at once: the result is the set written by the
// first sweep that fitted nothing, or by the last sweep
// (lsq_result_is_second picks it).
hipError_t launch_lsq_sweep_pingpong(const int16_t* cur_in, const double* ex_in,
                                     const double* ey_in, int16_t* cur_out, double* ex_out,
                                     double* ey_out, const int16_t* tgt, const double* gx,
                                     const double* gy, int n0, int n1, LsqWork w, int sweep,
                                     hipStream_t s);
hipError_t launch_fill_u32(unsigned* p, int n, unsigned v, hipStream_t s);
// After ping-pong sweeps 0..sweeps-1 over the first set (cur, ex, ey) and
// the second (w.cur2, w.ex2, w.ey2): when the last of them still fitted
// cells, one workgroup carries the reference's loop on over the listed cells
// (block barriers between sweeps) until a sweep fits nothing, and leaves the
// result in the set lsq_result_is_second names.  Its first sweep covers the
// list; later ones only the cells still unknown and those fitted by the
// sweep before (copied across, so each set ends up with every fit), kept in
// w.tail.  w.counters[sweeps] receives the sweeps it ran (0 when the loop
// had ended).  It ends within (list entries + 1) sweeps: every sweep but
// the last fills a listed cell.
hipError_t launch_lsq_sweep_tail(int16_t* cur, double* ex, double* ey, const int16_t* tgt,
                                 const double* gx, const double* gy, int n0, int n1, LsqWork w,
                                 int sweeps, hipStream_t s);

// True when ping-pong sweeps 0..sweeps-1 (sweep q writing the second set when
// q is even) left their result in the second buffer set.
__device__ __forceinline__ bool lsq_result_is_second(const unsigned* counters, int sweeps) {
  int last = 0;
  while (last < sweeps - 1 && counters[last] != 0u) ++last;
  return (last & 1) == 0;
}

}  // namespace axs
