// Drop-in for umappp/NeighborList.hpp (reference NeighborList.hpp:28-29).
// The reference aliases knncolle::NeighborList; this library has no knncolle
// dependency and spells the same type out: per observation, a vector of
// (neighbour index, distance or weight) pairs, sorted by increasing distance,
// unique, never containing the observation itself.
#ifndef UMAPPP_AMD_NEIGHBOR_LIST_HPP
#define UMAPPP_AMD_NEIGHBOR_LIST_HPP

#include <utility>
#include <vector>

namespace umappp {

template <typename Index_, typename Float_>
using NeighborList = std::vector<std::vector<std::pair<Index_, Float_>>>;

}  // namespace umappp

#endif
