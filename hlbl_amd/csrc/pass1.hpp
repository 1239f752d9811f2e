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
// d_units (T = float, optional): the (pair, flavour) units pair * 2 + f as
// `groups` lists of unit_slots entries (-1: none), list k run on the workgroup slots
// k, k + groups, ... that the dispatcher deals out in turn (with 8 lists, one list per
// XCD, so units of a list that share a block share its L2).
template <class T>
void launch_pass1_planes(const WinPair* d_pairs, const WinPair& single, int n_pairs,
                         const hlbl_geom& global, const hlbl_geom& local, const hlbl_coord& proc,
                         int flav_mask, T* part, double* red1, double* xi, hipStream_t st,
                         const int* d_units = nullptr, int unit_slots = 0, int groups = 1);

// The unit lists for d_units: the 2 n_pairs units of pairs[0..n) dealt into
// `groups` lists of balanced length (unit_slots = ceil(2 n / groups)), greedily by
// the fewest new (block, flavour) halves, then improved by swaps.
void pass1_group_units(const WinPair* pairs, int n_pairs, int* units, int unit_slots, int groups);

}  // namespace hlbl
