#include <LORCache.h>

namespace
{
constexpr std::uint32_t kDisabled = 0xFFFFFFFFu;
}

LORCache::LORCache(const ProjData& proj, int nSubsets)
  : mNSubsets(nSubsets),
    mNViewsPerSubset(0),
    mNCrystalsPerRing(proj.getHeader().nCrystalsPerRing),
    mSegOffset(proj.getGeometry().segOffset)
{
  proj.checkNSubsets(nSubsets);
  const auto& g = proj.getGeometry();
  const int nTang = proj.getHeader().nTangCoords;
  mNViewsPerSubset = g.nViews / nSubsets;
  mNBinsPerView.resize(2 * mSegOffset + 1);
  for (int seg = -mSegOffset; seg <= mSegOffset; ++seg)
    mNBinsPerView[seg + mSegOffset] = g.getNAxialCoords(seg) * nTang;
  mLors.assign(nSubsets, std::vector<std::vector<LOR>>(2 * mSegOffset + 1));
  for (int subset = 0; subset < nSubsets; ++subset)
    for (int seg = -mSegOffset; seg <= mSegOffset; ++seg)
    {
      auto& list = mLors[subset][seg + mSegOffset];
      list.reserve(static_cast<size_t>(mNViewsPerSubset) *
                   mNBinsPerView[seg + mSegOffset]);
      for (int sv = 0; sv < mNViewsPerSubset; ++sv)
      {
        const int view = sv * nSubsets + subset;
        for (int ax = 0; ax < g.getNAxialCoords(seg); ++ax)
        {
          const auto [s1, s2] = proj.getCrystalAxialCoord(seg, ax);
          for (int t = -g.tangCoordOffset; t < -g.tangCoordOffset + nTang; ++t)
          {
            const auto [a1, a2] = proj.getCrystalAngCoord(view, t);
            list.push_back({static_cast<std::uint32_t>(s1 * mNCrystalsPerRing + a1),
                            static_cast<std::uint32_t>(s2 * mNCrystalsPerRing + a2)});
          }
        }
      }
    }
}

int LORCache::setSubsetAndSegment(int subset, int segment)
{
  mCurrentSubset = subset;
  mCurrentSegment = segment;
  return static_cast<int>(mLors[subset][segment + mSegOffset].size());
}

std::tuple<bool, int, int, int, int, int> LORCache::getLOR(int index) const
{
  const int perView = mNBinsPerView[mCurrentSegment + mSegOffset];
  const int sv = index / perView;
  const int rem = index - sv * perView;
  const int bin = (sv * mNSubsets + mCurrentSubset) * perView + rem;
  const LOR& lor = mLors[mCurrentSubset][mCurrentSegment + mSegOffset][index];
  if (lor.crystal1 == kDisabled) return {false, bin, 0, 0, 0, 0};
  const int c1 = static_cast<int>(lor.crystal1);
  const int c2 = static_cast<int>(lor.crystal2);
  return {true, bin, c1 / mNCrystalsPerRing, c1 % mNCrystalsPerRing,
          c2 / mNCrystalsPerRing, c2 % mNCrystalsPerRing};
}

void LORCache::disableLOR(int index)
{
  mLors[mCurrentSubset][mCurrentSegment + mSegOffset][index].crystal1 =
    kDisabled;
}
