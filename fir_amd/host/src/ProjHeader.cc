#include <ProjHeader.h>
#include <console.h>

void ProjHeader::setDefaults()
{
  nRings = 0;
  nCrystalsPerRing = 0;
  segmentSpan = 1;
  nSegments = 1;
  nTangCoords = 0;
}

void ProjHeader::check()
{
  if (nRings <= 0) error("\"number of rings\" must be provided and > 0");
  if (nCrystalsPerRing <= 0)
    error("\"number of crystals per ring\" must be provided and > 0");
  if (nCrystalsPerRing % 4 != 0)
    error("\"number of crystals per ring\" must be a multiple of four");
  if (segmentSpan <= 0) error("\"segment span\" must be > 0");
  if (segmentSpan % 2 != 1) error("\"segment span\" must be an odd number");
  const int maxSpan = 2 * nRings - 1;
  if (segmentSpan > maxSpan)
    error("For ", nRings, " ring(s), \"segment span\" must not be > ",
          maxSpan);
  if (nSegments < 0) error("\"number of segments\" must not be negative");
  if (nSegments % 2 != 1) error("\"number of segments\" must be an odd number");
  int maxSeg = maxSpan / segmentSpan;
  if (maxSeg % 2 == 0) --maxSeg;
  if (nSegments > maxSeg)
    error("For ", nRings, " ring(s) and a segment span of ", segmentSpan,
          ", \"number of segments\" must not be > ", maxSeg);
  if (nTangCoords < 0)
    error("\"number of tangential coordinates\" must not be negative");
  const int maxTang = nCrystalsPerRing - 1;
  if (nTangCoords == 0)
    nTangCoords = maxTang;
  else if (nTangCoords > maxTang)
    error("For ", nCrystalsPerRing,
          " crystals per ring, \"number of tangential coordinates\" must not "
          "be > ",
          maxTang);
}

void ProjGeometry::fill(const ProjHeader& h)
{
  segOffset = (h.nSegments - 1) / 2;
  tangCoordOffset = h.nTangCoords / 2;
  nAxialCoords.assign(h.nSegments, 0);
  for (int seg = -segOffset; seg <= segOffset; ++seg)
  {
    const int a = seg < 0 ? -seg : seg;
    int n;
    if (h.segmentSpan == 1)
      n = h.nRings - a;
    else
    {
      // central segment spans every slice; each step away loses the
      // diagonals it shares with its neighbour
      n = 2 * h.nRings - 1;
      if (a >= 1) n -= h.segmentSpan + 1;
      if (a >= 2) n -= 2 * h.segmentSpan * (a - 1);
    }
    nAxialCoords[seg + segOffset] = n;
  }
  nViews = h.nCrystalsPerRing / 2;
  nBins = 0;
  for (int n : nAxialCoords) nBins += n * nViews * h.nTangCoords;
  halfSegmentSpan = (h.segmentSpan - 1) / 2;
  if (h.segmentSpan == 1)
    maxRingDiff = segOffset;
  else
    maxRingDiff = halfSegmentSpan + (segOffset > 0 ? segOffset * h.segmentSpan : 0);
}
