// Opposite-direction pairs of a lattice, shared by host and device code.
// Pair k couples direction lo(k) with opp(lo(k)); lo runs over the directions
// whose opposite has a larger index, in ascending order, so for the 3D sets
// lo(k) = 2k + 1 and for D2Q9 lo = {1, 2, 5, 6}.
#pragma once
#include "lattice_tables.h"

template <int Q>
struct Pairs {
  static constexpr int NP = (Q - 1) / 2;
  static constexpr int lo(int k) {
    int c = 0;
    for (int i = 1; i < Q; i++) {
      if (Lattice<Q>::opp[i] > i) {
        if (c == k) return i;
        c++;
      }
    }
    return -1;
  }
};
