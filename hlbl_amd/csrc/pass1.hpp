// Pass 1 (weighted site sums of the spin-pair colour products) for one pair or
// a whole window; host-side launcher.
#pragma once

#include <hip/hip_runtime.h>

#include "window.hpp"

namespace hlbl {

// buffer sizes: plane partials (elements of T) and first-level sums (doubles)
size_t pass1_part_len(const hlbl_geom& local, int n_pairs);
size_t pass1_red_len(const hlbl_geom& local, int n_pairs);

// xi[pair][f][p][w][144][2] (stride 2*8*5*144*2 per pair) of the pairs d_pairs[0..n)
// (device array), or of `single` when d_pairs is null (n_pairs = 1). Flavours
// outside flav_mask get zeros. T = double: FP64 products; T = float: FP32
// products and plane sums. The reductions are FP64 Kahan in a fixed order.
template <class T>
void launch_pass1_planes(const WinPair* d_pairs, const WinPair& single, int n_pairs,
                         const hlbl_geom& global, const hlbl_geom& local, const hlbl_coord& proc,
                         int flav_mask, T* part, double* red1, double* xi, hipStream_t st);

}  // namespace hlbl
