// LORCache.h -- per-(subset, segment) list of crystal pairs, the
// reference's LORCache interface (LORCache.h:5-47). Order within a
// (subset, segment): (subview, axial, tang) with view = subview * nSubsets
// + subset. Crystal ids are 32-bit (slice * nCrystalsPerRing + angle), so
// long-axial scanners no longer overflow the reference's 16-bit ids.
// The GPU kernels enumerate the same LORs arithmetically; this class keeps
// the host-side API for callers that walk LORs themselves.
#pragma once

#include <ProjData.h>

#include <cstdint>
#include <tuple>
#include <vector>

struct LOR
{
  std::uint32_t crystal1; // 0xFFFFFFFF: disabled
  std::uint32_t crystal2;
};

class LORCache
{
public:
  LORCache(const ProjData& proj, int nSubsets);

  // Select (subset, segment); returns its number of LORs.
  int setSubsetAndSegment(int subset, int segment);

  // {valid, bin index within the segment, crystalAxialCoord1,
  //  crystalAngCoord1, crystalAxialCoord2, crystalAngCoord2}
  std::tuple<bool, int, int, int, int, int> getLOR(int index) const;

  void disableLOR(int index);

private:
  int mNSubsets, mNViewsPerSubset, mNCrystalsPerRing, mSegOffset;
  std::vector<int> mNBinsPerView;                 // [segment]
  std::vector<std::vector<std::vector<LOR>>> mLors; // [subset][segment][lor]
  int mCurrentSubset = 0, mCurrentSegment = 0;
};
