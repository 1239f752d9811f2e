// The window-batched path: per-window plan, buffers and the captured graph.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <tuple>
#include <vector>

#include "context.hpp"

namespace hlbl {

// below this distance (muon units) from the origin, |x| or |x - y|, the
// mixed-precision forms evaluate the QED kernel in FP64 (Lambda subtraction
// cancels there)
constexpr double kNearRadius = 0.25;

// one pair of a window as the kernels read it
struct WinPair {
  const double* S;   // fwd_src: block of x
  const double* Y;   // fwd_y: block of y
  hlbl_coord gsx;
  hlbl_coord yv;
  int yv_slot;       // the slot of yv's coefficient field in the field store
  int pad;
};

// A window call reads one coefficient field per distinct yv from the contractor's field
// store and builds the ones the store lacks: n_unique of them, listed in d_yvs with
// their store slots in d_slots.
struct WindowPlan {
  int n_pairs = 0, n_unique = 0, form = 0, n_samples = 0;
  double xunit = 0.0;
  DevBuf<WinPair> d_pairs;
  DevBuf<hlbl_coord> d_yvs;  // the yv whose fields this call builds
  DevBuf<int> d_slots;       // their slots in the store
  PinBuf<WinPair> h_pairs;
  PinBuf<hlbl_coord> h_yvs;
  PinBuf<int> h_slots;
  std::vector<int> call_slots;  // the store slots of the call's distinct yv
  float* coef = nullptr;        // the field store [slot][V][640]
  int batch = 0;            // pass-1 units per list
  DevBuf<int> d_order;      // pass-1 unit order (pair * 2 + f), lists of `batch` that share blocks
  PinBuf<int> h_order;
  DevBuf<float> part1;     // pass-1 plane partials
  DevBuf<double> red1, xi, dz, gdz;
  DevBuf<unsigned short> bsplit;  // pass-2 B operands, FP16 hi/lo in MFMA fragment order
  DevBuf<int> bexp;               // their column exponents [pair][f][160]
  DevBuf<float> ax;        // L0(x, 0) and L0(0, x) of every site [2][V][384]
  DevBuf<double> cst64;    // pair constants L0(0,y), L0(y,0) per distinct yv [n_unique][768]
  DevBuf<float> cst32;
  DevBuf<float> chbuf;     // surrogate channels [2 + 3 n_unique][V][32]
  DevBuf<double> part2, result;
  PinBuf<double> h_result;
  DevBuf<size_t> d_samples;
  DevBuf<double> judge;    // [n_unique][n_samples] of the fields built
  PinBuf<double> h_judge;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~WindowPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

// kernels of a window on `st`; stage 0 = pass 1 and its reductions (recorded into the
// window's graph), 1 = the coefficient fields the store lacks (and their judge), 2 =
// pass 2 and its reduction (recorded)
void window_stage(hlbl_contractor* c, WindowPlan& w, int stage, hipStream_t st);
size_t window_part2_len(const hlbl_contractor* c, int n_pairs);
size_t window_bsplit_len(int n_pairs);  // FP16 elements of the pass-2 B operands
size_t window_bexp_len(int n_pairs);    // their column exponents

// dz and gdz (FP64) of n_pairs reduced pass-1 sums xi[pair][...]
void launch_forms_batched(const double* xi, double* dz, double* gdz, int n_pairs, hipStream_t st);

// surrogate channels of the six site evaluations, for every site and yv
void launch_surrogate_channels(hlbl_contractor* c, WindowPlan& w, hipStream_t st);

}  // namespace hlbl
