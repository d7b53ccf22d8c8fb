// ProjHeader.h -- sinogram header and derived geometry (reference
// ProjHeader.h:15-80). Bins are indexed by segment, view, axial and
// tangential coordinates; the axial count depends on the segment.
#pragma once

#include <vector>

struct ProjHeader
{
  int nRings;
  int nCrystalsPerRing;
  int segmentSpan;
  int nSegments;
  int nTangCoords; // 0 selects the maximum, nCrystalsPerRing - 1

  bool operator==(const ProjHeader& o) const
  {
    return nRings == o.nRings && nCrystalsPerRing == o.nCrystalsPerRing &&
      segmentSpan == o.segmentSpan && nSegments == o.nSegments &&
      nTangCoords == o.nTangCoords;
  }
  bool operator!=(const ProjHeader& o) const { return !(*this == o); }

  void setDefaults();
  void check();
};

struct ProjGeometry
{
  int nBins;
  int segOffset;
  int tangCoordOffset;
  std::vector<int> nAxialCoords;
  int nViews;
  int halfSegmentSpan;
  int maxRingDiff;

  int getNAxialCoords(int seg) const { return nAxialCoords[seg + segOffset]; }
  void fill(const ProjHeader& header);
};
