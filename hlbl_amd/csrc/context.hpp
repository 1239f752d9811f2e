// Library state behind the opaque handles of the C ABI, and error plumbing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hlbl_amd.h"
#include "qed.hpp"

namespace hlbl {

// thrown inside the library, turned into a status code at the C ABI
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_last_error(const std::string& msg);

inline void check_hip(hipError_t err, const char* what) {
  if (err != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(err));
}

// run f, map exceptions to -1 + last error
template <class F>
int guarded(F&& f) {
  try {
    f();
    set_last_error("");
    return 0;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return -1;
  }
}

// an owning device allocation
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void ensure(size_t count) {
    if (count <= n) return;
    release();
    check_hip(hipMalloc(&p, count * sizeof(T)), "hipMalloc");
    n = count;
  }
};

template <class T>
struct PinBuf {
  T* p = nullptr;
  size_t n = 0;
  PinBuf() = default;
  PinBuf(const PinBuf&) = delete;
  PinBuf& operator=(const PinBuf&) = delete;
  ~PinBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  void ensure(size_t count) {
    if (count <= n) return;
    release();
    check_hip(hipHostMalloc(&p, count * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    n = count;
  }
};

// MLP surrogate sizes
constexpr int kMlpIn = 3;
constexpr int kMlpHidden = 128;
constexpr int kMlpLayers = 4;  // hidden layers
constexpr int kMlpOut = qed::kNChUsed;
// parameter blocks: W0[128][3] b0[128], W1..W3[128][128] b[128], W4[42][128] b4[42]
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden +
                           (kMlpLayers - 1) * (kMlpHidden * kMlpHidden + kMlpHidden) +
                           kMlpOut * kMlpHidden + kMlpOut;

}  // namespace hlbl

struct hlbl_qed {
  int device = 0;
  hlbl::DevBuf<double> table64;  // [NA][NC][NB][NCh]
  hlbl::DevBuf<float> table32;
  std::vector<double> host_table;
  // MLP: FP32 master parameters, BF16 copy, output scale/offset per channel
  hlbl::DevBuf<float> mlp_w;            // FP32 master parameters
  hlbl::DevBuf<float> mlp_wq;           // the same rounded to BF16, as floats
  hlbl::DevBuf<unsigned short> mlp_w_bf16;
  hlbl::DevBuf<float> mlp_scale;  // [2][kMlpOut]: offset, scale (normalised -> physical)
  bool mlp_ready = false;
  uint64_t mlp_gen = 0;  // bumped whenever the network or its domain changes
  double mlp_radius = hlbl::qed::kRadius;  // trained domain: |x|, |y| <= mlp_radius (muon units)
  // and |x|, |y| >= mlp_inner: below it (the rows L0(x, 0), L0(0, x) among them, where the
  // weight functions bend like |y| log |y|) the table serves
  double mlp_inner = 0.125;
  double* noalloc_scratch = nullptr;  // this device's fixed scratch array (qed_noalloc.hip)
};

namespace hlbl {
double* noalloc_scratch_setup(int device);
}  // namespace hlbl

namespace hlbl {
struct WindowCache;
void window_cache_free(WindowCache* w);
}  // namespace hlbl

namespace hlbl {
inline hlbl_tuning default_tuning() { return hlbl_tuning{32, 8, 2, 4, 1, 1, 2, 2}; }
}  // namespace hlbl

struct hlbl_contractor {
  hlbl_geom global{}, local{};
  hlbl_tuning tune = hlbl::default_tuning();
  hlbl_coord proc{};
  hlbl_qed* qed = nullptr;
  size_t volume = 0;
  int n_cu = 0;
  int nwg2 = 0;
  hlbl::DevBuf<double> part1;   // pass-1 plane partials [2][planes][3][2304]
  hlbl::DevBuf<double> red1;    // pass-1 chunk sums [2][chunks][5][2304]
  hlbl::DevBuf<double> xi;      // reduced [2][8][5][144][2]
  hlbl::DevBuf<double> dz;      // dzu, dzsu of both flavours [2][10][144][2]
  hlbl::DevBuf<double> gdz;     // g_dzu, g_dzsu of both flavours [2][40][144][2]
  hlbl::DevBuf<double> bmat;    // pass-2 B operands [2][288][48]
  hlbl::DevBuf<double> coef;    // pass-2 coefficient field [V][16][40]
  hlbl::DevBuf<double> kcst;    // the pair constants L0(0,y), L0(y,0) [2][384]
  hlbl::DevBuf<double> ax64;    // FP64 site-only rows L0(x,0), L0(0,x) at relative sites [2][V][384]
  double ax64_xunit = -1.0;     //   for this xunit (none: negative)
  hlbl::DevBuf<double> part2;   // pass-2 partials [nwg2][12]
  hlbl::DevBuf<double> pi2p2;   // the 2+2 pieces' Pi_{mu nu}(z) [V][16]
  hlbl::DevBuf<double> part234; // P2..P4 workgroup partials [workgroups][320]
  hlbl::DevBuf<double> p1buf;   // the 2+2 step of a run: P1 [4][4][4][Lmax]
  hlbl::DevBuf<double> p234buf; //   and P2..P4 [targets][320]
  hlbl::DevBuf<double> result;  // [2]
  hlbl::PinBuf<double> host_result;
  hlbl::WindowCache* window = nullptr;
  ~hlbl_contractor() { hlbl::window_cache_free(window); }
};
