// reconAlgos.h -- OSEM of the FIR host API (reference reconAlgos.h:9-45),
// run on the GPU with every image and sinogram resident in HBM.
#pragma once

#include <ProjData.h>
#include <ScannerData.h>
#include <VolData.h>

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
};

namespace reconAlgos
{
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

// Resolution-modelled OSEM as the reference computes it (reference
// reconAlgos.cc:203-367): the sensitivity frames are convolved once (in
// place), and each sub-iteration convolves the backprojection, then
// x /= sensitivity frame, x *= bp, Gaussian every convolutionInterval
// sub-iterations, circular cut. The forward projection uses x itself, as
// the reference's does.
// OSEM (reso false) or OSEM_ResoReco (reso true, in.sens convolved in
// place) on device-resident inputs; inputProj gives the sinogram geometry
// only (its bins are not read). An extension of the reference API.
void OSEMOnDevice(const ProjData& inputProj, const ScannerData& scanner,
                  VolData& outputVol, const std::string& outputVolFileName,
                  const OSEMCoreParams& params, const DeviceOSEMInputs& in,
                  bool reso = false);

void OSEM_ResoReco(const ProjData& inputProj, const ScannerData& scanner,
                   VolData& outputVol, const std::string& outputVolFileName,
                   const OSEMCoreParams& params, VolData& sensitivityMap,
                   const std::optional<ProjData>& biasProj);
}
