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
// d_order (T = float, optional): the (pair, flavour) units pair * 2 + f in the order
// an XCD runs them on each of its planes (the planes are dealt to the XCDs in turn);
// null: unit order.
template <class T>
void launch_pass1_planes(const WinPair* d_pairs, const WinPair& single, int n_pairs,
                         const hlbl_geom& global, const hlbl_geom& local, const hlbl_coord& proc,
                         int flav_mask, T* part, double* red1, double* xi, hipStream_t st,
                         const int* d_order = nullptr);

// The order for d_order: the 2 n_pairs units of pairs[0..n) dealt into lists of at
// most `batch` units (balanced; ceil(2 n / batch) lists), greedily by the fewest new
// (block, flavour) halves, then improved by swaps; the lists one after another.
void pass1_order_units(const WinPair* pairs, int n_pairs, int batch, int* order);

}  // namespace hlbl
