// The QED kernel at points with no allocation: hlbl_qed_kernel's plain form with its
// per-point scratch (three L0 evaluations of 384 doubles per argument pair) in a fixed
// device array of this library instead of a per-call allocation, for callers that
// cannot allocate (code under graph capture, tight serving loops, harnesses that
// forbid allocation).
//
// The array is one per GPU (its own copy, at its own address, held until the process
// ends). Every call on a GPU shares it: calls that overlap on one GPU (two streams, or
// two host threads) race on it and must be ordered by the caller. Nothing in it needs
// to start at zero: each point's three L0 evaluations are written before they are read.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/hlbl_amd.h"
#include "context.hpp"
#include "contract_kernels.hpp"
#include "qed.hpp"

namespace {

// at most kNoAllocPoints argument pairs per call: 2048 x 3 x 384 doubles, 18.9 MB per GPU
constexpr int kNoAllocPoints = 2048;
constexpr size_t kScratchLen = (size_t)kNoAllocPoints * 3 * hlbl::qed::kNL;

}  // namespace

static __device__ __align__(256) double scratch[kScratchLen];

namespace hlbl {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch[kMaxDevices] = {};

}  // namespace

// The device address of the scratch array on `device`, which must be the current
// device: looked up once per device (hipGetSymbolAddress), then kept. The array's
// name is a host-side stand-in, not a device address; only this pointer is passed on.
double* noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error("noalloc scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[device]) {
    void* p = nullptr;
    check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch[device] = static_cast<double*>(p);
  }
  return g_scratch[device];
}

}  // namespace hlbl

using namespace hlbl;

extern "C" int hlbl_qed_kernel_noalloc(hlbl_qed* q, const double* d_x, const double* d_y, int n,
                                       double* d_kerv, hlbl_stream stream) {
  return guarded([&] {
    if (!q) throw Error("qed is NULL");
    if (n < 0) throw Error("n < 0");
    if (n > kNoAllocPoints)
      throw Error("hlbl_qed_kernel_noalloc: at most 2048 argument pairs per call");
    if (n == 0) return;
    if (!d_x || !d_y || !d_kerv) throw Error("hlbl_qed_kernel_noalloc: NULL device pointer");
    if (!q->noalloc_scratch) throw Error("hlbl_qed_kernel_noalloc: scratch not set up");
    launch_qed_points(q->table64.p, d_x, d_y, n, d_kerv, q->noalloc_scratch,
                      reinterpret_cast<hipStream_t>(stream));
  });
}
