// reconAlgos.h -- OSEM of the FIR host API (reference reconAlgos.h:9-45),
// run on the GPU with every image and sinogram resident in HBM.
#pragma once

#include <ProjData.h>
#include <ScannerData.h>
#include <VolData.h>

#include <memory>
#include <optional>
#include <string>
#include <vector>

struct OSEMCoreParams
{
  int nIterations{1};
  int nSubsets{1};
  int saveInterval{0};
  float cutRadius{0.0};
  int convolutionInterval{0};
  std::vector<float> fwhmXYZ{0.0, 0.0, 0.0};
};

// Device-resident data of one OSEM run: device pointers (nBins floats for
// y and bias, nSubsets frames for sens). No counterpart in the reference
// API: FIR_OSEM keeps its inputs in HBM from the moment they are read.
struct DeviceOSEMInputs
{
  const float* y = nullptr;    // measured sinogram (attenuation applied)
  const float* bias = nullptr; // additive term, or none
  float* sens = nullptr;       // sensitivity frames
  // fill sens first with the sensitivity of the projector the iterations
  // use (the path cache's for its subsets, else the walk's), copying it to
  // sensOut (host frames) when given
  bool computeSens = false;
  VolData* sensOut = nullptr;
  // a path cache built beforehand for this grid and subset count (not
  // owned); without one, the run builds its own when FIR_PATH_CACHE asks
  fir_cache* cache = nullptr;
};

namespace reconAlgos
{
// The path cache FIR_PATH_CACHE asks for (snorm16 | fp16 | fp32; none when
// it is unset, empty or "off"), for OSEM over grid v with nSubsets subsets:
// whole, or partial when FIR_PATH_CACHE_SUBSETS / FIR_PATH_CACHE_GB cap it
// or the whole cache does not fit. It depends on the geometry only (the
// device context's scanner and sinogram tables), so FIR_OSEM builds it on a
// second host thread while it reads the data. No counterpart in the
// reference.
struct PathCacheHandle
{
  fir_cache* cache = nullptr;
  int nCached = 0; // subsets 0 .. nCached-1 are held
  PathCacheHandle() = default;
  PathCacheHandle(const PathCacheHandle&) = delete;
  PathCacheHandle& operator=(const PathCacheHandle&) = delete;
  ~PathCacheHandle();
};
bool pathCacheRequested();
std::unique_ptr<PathCacheHandle> buildPathCache(const fir_vol_desc& v,
                                                int nSubsets);

// OSEM on the active frame of outputVol (the initial estimate). Per
// sub-iteration: projector pass over the subset's LORs (bias added to the
// forward projection when given), bp /= sensitivity frame, x *= bp,
// Gaussian every convolutionInterval sub-iterations, circular cut.
// Intermediate images go to <outputVolFileName>_subiter_<n> every
// saveInterval sub-iterations.
void OSEM(const ProjData& inputProj, const ScannerData& scanner,
          VolData& outputVol, const std::string& outputVolFileName,
          const OSEMCoreParams& params, const VolData& sensitivityMap,
          const std::optional<ProjData>& biasProj);

// OSEM (reso false) or OSEM_ResoReco (reso true, in.sens convolved in
// place) on device-resident inputs; inputProj gives the sinogram geometry
// only (its bins are not read). An extension of the reference API.
void OSEMOnDevice(const ProjData& inputProj, const ScannerData& scanner,
                  VolData& outputVol, const std::string& outputVolFileName,
                  const OSEMCoreParams& params, const DeviceOSEMInputs& in,
                  bool reso = false);

// Resolution-modelled OSEM as the reference computes it (reference
// reconAlgos.cc:203-367): the sensitivity frames are convolved once (in
// place), and each sub-iteration convolves the backprojection, then
// x /= sensitivity frame, x *= bp, Gaussian every convolutionInterval
// sub-iterations, circular cut. The forward projection uses x itself, as
// the reference's does.
void OSEM_ResoReco(const ProjData& inputProj, const ScannerData& scanner,
                   VolData& outputVol, const std::string& outputVolFileName,
                   const OSEMCoreParams& params, VolData& sensitivityMap,
                   const std::optional<ProjData>& biasProj);
}
