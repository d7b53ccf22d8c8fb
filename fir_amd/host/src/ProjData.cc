#include <ProjData.h>
#include <ProjInterfileReader.h>
#include <console.h>
#include <device.h>

#include <algorithm>

ProjData::ProjData() = default;

ProjData::ProjData(const std::string& file, ConstructionMode mode,
                   types::BinValue init)
{
  read(file, mode, init);
}

ProjData::ProjData(const ProjData& proj, ConstructionMode mode,
                   types::BinValue init)
{
  copy(proj, mode, init);
}

ProjData::ProjData(ProjData&& o) noexcept { *this = std::move(o); }

ProjData& ProjData::operator=(ProjData&& o) noexcept
{
  if (this == &o) return *this;
  mHeader = o.mHeader;
  mGeometry = o.mGeometry;
  mStorage = std::move(o.mStorage);
  mSegments = std::move(o.mSegments);
  o.mStorage.clear();
  o.mSegments.clear();
  return *this;
}

void ProjData::allocate(bool initialize, types::BinValue init)
{
  mStorage.assign(static_cast<size_t>(mGeometry.nBins),
                  initialize ? init : 0.0f);
  mSegments.clear();
  size_t off = 0;
  for (int seg = -mGeometry.segOffset; seg <= mGeometry.segOffset; ++seg)
  {
    mSegments.push_back(mStorage.data() + off);
    off += static_cast<size_t>(mGeometry.nViews) *
      mGeometry.getNAxialCoords(seg) * mHeader.nTangCoords;
  }
}

void ProjData::read(const std::string& file, ConstructionMode mode,
                    types::BinValue init)
{
  ProjInterfileReader reader(file);
  mHeader = reader.getHeader();
  mGeometry = reader.getGeometry();
  switch (mode)
  {
  case ConstructionMode::ALLOCATE:
    allocate(false, 0.0f);
    break;
  case ConstructionMode::INITIALIZE:
    allocate(true, init);
    break;
  case ConstructionMode::READ_DATA:
    allocate(false, 0.0f);
    reader.readData(getDataArray());
    break;
  }
}

void ProjData::copy(const ProjData& proj, ConstructionMode mode,
                    types::BinValue init)
{
  mHeader = proj.mHeader;
  mGeometry = proj.mGeometry;
  switch (mode)
  {
  case ConstructionMode::ALLOCATE:
    allocate(false, 0.0f);
    break;
  case ConstructionMode::INITIALIZE:
    allocate(true, init);
    break;
  case ConstructionMode::READ_DATA:
    allocate(false, 0.0f);
    std::copy(proj.mStorage.begin(), proj.mStorage.end(), mStorage.begin());
    break;
  }
}

void ProjData::write(const std::string& file) const
{
  if (mStorage.empty()) error("Projection data is not allocated");
  ProjInterfileReader::writeProjInterfile(file, mHeader, getDataArray());
}

void ProjData::printContent() const
{
  echo("= Projection header:");
  printEmptyLine();
  echo("== Scanner basic geometry:");
  printValue("number of rings", mHeader.nRings);
  printValue("number of crystals per ring", mHeader.nCrystalsPerRing);
  printEmptyLine();
  echo("== Michelogram compression:");
  printValue("segment span", mHeader.segmentSpan);
  printEmptyLine();
  echo("== Projection dimensions:");
  printValue("number of segments", mHeader.nSegments);
  printVector("number of axial coords", mGeometry.nAxialCoords);
  printValue("number of views", mGeometry.nViews);
  printValue("number of tangential coordinates", mHeader.nTangCoords);
  printEmptyLine();
  echo("== Others:");
  printValue("total number of bins", mGeometry.nBins);
  printValue("segment offset", mGeometry.segOffset);
  printValue("tangential coordinate offset", mGeometry.tangCoordOffset);
  printValue("half the segment span", mGeometry.halfSegmentSpan);
  printValue("maximum ring difference", mGeometry.maxRingDiff);
  printEmptyLine();
}

void ProjData::checkNSubsets(int nSubsets) const
{
  if (nSubsets <= 0 || mGeometry.nViews % nSubsets != 0)
    error("Number of subsets must be a divisor of ", mGeometry.nViews);
}

void ProjData::setAllBins(types::BinValue value)
{
  if (mStorage.empty()) error("Projection not allocated");
  std::fill(mStorage.begin(), mStorage.end(), value);
}

std::tuple<int, int> ProjData::getCrystalAxialCoord(int seg, int ax) const
{
  if (mHeader.segmentSpan == 1)
  {
    // span 1: segment = ring difference, axial = lower ring
    if (seg == 0) return {2 * ax, 2 * ax};
    if (seg > 0) return {2 * (ax + seg), 2 * ax};
    return {2 * ax, 2 * (ax - seg)};
  }
  if (seg == 0) return {ax, ax};
  const int aseg = seg < 0 ? -seg : seg;
  const int mid = aseg * mHeader.segmentSpan;
  const int half = mGeometry.halfSegmentSpan;
  if (ax < half)
  {
    const int rd = mid - half + ax;
    return seg > 0 ? std::tuple<int, int>{2 * rd, 0}
                   : std::tuple<int, int>{0, 2 * rd};
  }
  const int inv = mGeometry.getNAxialCoords(seg) - ax - 1;
  if (inv < half)
  {
    const int rd = mid - half + inv;
    const int top = 2 * mHeader.nRings - 2;
    return seg > 0 ? std::tuple<int, int>{top, top - 2 * rd}
                   : std::tuple<int, int>{top - 2 * rd, top};
  }
  const int adj = ax - half;
  return seg > 0 ? std::tuple<int, int>{2 * mid + adj, adj}
                 : std::tuple<int, int>{adj, 2 * mid + adj};
}

std::tuple<int, int> ProjData::getCrystalAngCoord(int view, int t) const
{
  // t/2 rounded toward -inf for odd negative t, plus the odd-bin shift
  const int shift = t >= 0 ? t / 2 : (t - 1) / 2;
  const int odd = (t < 0 ? -t : t) % 2;
  const int n = mHeader.nCrystalsPerRing;
  int a = view + shift + odd;
  int b = view + mGeometry.nViews - shift;
  if (a < 0) a += n;
  else if (a >= n) a -= n;
  if (b < 0) b += n;
  else if (b >= n) b -= n;
  return {a, b};
}

bool ProjData::getBinCoordinates(int r1, int c1, int r2, int c2, int* seg,
                                 int* view, int* axialCoord,
                                 int* tangCoord) const
{
  const int n = mHeader.nCrystalsPerRing;
  const int d = r1 > r2 ? r1 - r2 : r2 - r1;
  if (d > mGeometry.maxRingDiff) return false;
  // tangential: signed distance of the crystal separation from n/2, the
  // sign given by the half-turn the crystal sum lies in
  const int sum = c1 + c2;
  const int s1 = (sum >= n / 2 && sum < 3 * n / 2) ? -1 : 1;
  const int t = s1 * ((c2 > c1 ? c2 - c1 : c1 - c2) - n / 2);
  if (t < -mGeometry.tangCoordOffset ||
      t >= mHeader.nTangCoords - mGeometry.tangCoordOffset)
    return false;
  // axial: lower ring (span 1), or ring sum minus the first ring sum of
  // the fused segment
  int m, ax;
  const int half = mGeometry.halfSegmentSpan;
  if (mHeader.segmentSpan == 1)
  {
    m = d;
    ax = (r1 + r2 - d) / 2;
  }
  else
  {
    m = d <= half ? 0 : 1 + (d - half - 1) / mHeader.segmentSpan;
    ax = r1 + r2 - (m == 0 ? 0 : 1 + half + (m - 1) * mHeader.segmentSpan);
  }
  int s = m;
  if (m != 0)
  {
    // sign: both crystals moved to the parallel LOR through the centre
    // (tangential 0 or 1), compared, and the rings compared
    const int s2 = c1 < c2 ? 1 : -1;
    const int u = (t % 2 == 0) ? -t / 2 : -(t - 1) / 2;
    int k1 = c1 - s1 * s2 * u;
    int k2 = c2 + s1 * s2 * u;
    k1 += k1 >= n ? -n : (k1 < 0 ? n : 0);
    k2 += k2 >= n ? -n : (k2 < 0 ? n : 0);
    s = m * (r1 < r2 ? -1 : 1) * (k1 < k2 ? 1 : -1);
  }
  *seg = s;
  *view = ((sum + n / 2) % n) / 2;
  *axialCoord = ax;
  *tangCoord = t;
  return true;
}

void ProjData::rebinWeight()
{
  if (mStorage.empty()) error("Projection not allocated");
  device::rebinWeight(mHeader, mStorage.data(),
                      static_cast<long long>(mStorage.size()));
}

ProjData& ProjData::operator*=(const ProjData& in)
{
  if (mHeader != in.mHeader) error("Projections do not have the same size");
  device::clampedOp(device::ClampedOp::Mul, mStorage.data(),
                    in.mStorage.data(), static_cast<long long>(mStorage.size()));
  return *this;
}

void ProjData::exponential()
{
  device::expClamped(mStorage.data(), static_cast<long long>(mStorage.size()));
}

long long ProjData::index(int seg, int view, int ax, int t) const
{
  const int na = mGeometry.getNAxialCoords(seg);
  return (static_cast<long long>(view) * na + ax) * mHeader.nTangCoords + t +
    mGeometry.tangCoordOffset;
}

void ProjData::setBin(int seg, int view, int ax, int t, types::BinValue v)
{
  mSegments[seg + mGeometry.segOffset][index(seg, view, ax, t)] = v;
}

types::BinValue ProjData::getBin(int seg, int view, int ax, int t) const
{
  return mSegments[seg + mGeometry.segOffset][index(seg, view, ax, t)];
}

void ProjData::incrementBin(int seg, int view, int ax, int t)
{
  mSegments[seg + mGeometry.segOffset][index(seg, view, ax, t)] += 1.0f;
}

void ProjData::weightBin(int seg, int view, int ax, int t, types::BinValue w)
{
  mSegments[seg + mGeometry.segOffset][index(seg, view, ax, t)] *= w;
}
