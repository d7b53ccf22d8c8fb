#include <console.h>
#include <device.h>
#include <reconAlgos.h>

#include <string>

namespace reconAlgos
{
void OSEM(const ProjData& inputProj, const ScannerData& scanner,
          VolData& outputVol, const std::string& outputVolFileName,
          const OSEMCoreParams& p, const VolData& sens,
          const std::optional<ProjData>& biasProj)
{
  echo("OSEM:");
  scanner.checkProjData(inputProj);
  inputProj.checkNSubsets(p.nSubsets);
  if (sens.getNFrames() != p.nSubsets)
    error("Wrong number of frames (expected ", p.nSubsets, ")");
  if (sens.getHeader() != outputVol.getHeader())
    error("Sensitivity volume provided doesn't fit with output volume "
          "provided");
  if (biasProj && biasProj->getHeader() != inputProj.getHeader())
    error("Bias projection doesn't fit with input projection");
  if (p.fwhmXYZ.size() != 3)
    error("\"convolution FHWM XYZ in mm\" must contain three elements");

  const bool convolveFlag = p.convolutionInterval > 0 && p.fwhmXYZ[0] > 0.0 &&
    p.fwhmXYZ[1] > 0.0 && p.fwhmXYZ[2] > 0.0;

  device::useGeometry(scanner, inputProj.getHeader());
  fir_ctx* ctx = device::context();
  const fir_vol_desc v = device::describe(outputVol.getHeader());
  const size_t nv = static_cast<size_t>(outputVol.getNVoxelsPerFrame());
  const size_t nb = static_cast<size_t>(inputProj.getGeometry().nBins);
  const size_t vb = nv * sizeof(float);

  device::Buffer x(vb), bp(vb), y(nb * sizeof(float));
  device::Buffer s(vb * p.nSubsets), bias;
  x.upload(outputVol.getDataArray(), vb);
  y.upload(inputProj.getContiguousData(), nb * sizeof(float));
  s.upload(sens.getAllFrames(), vb * p.nSubsets);
  if (biasProj)
  {
    bias = device::Buffer(nb * sizeof(float));
    bias.upload(biasProj->getContiguousData(), nb * sizeof(float));
  }
  device::check(fir_fill(ctx, bp.f(), static_cast<int64_t>(nv), 0.0f));
  device::check(fir_cut_circle(ctx, &v, x.f(), 1, p.cutRadius));

  const int total = p.nIterations * p.nSubsets;
  for (int iter = 0; iter < p.nIterations; ++iter)
  {
    print("Iteration ", iter + 1, " of ", p.nIterations);
    for (int subset = 0; subset < p.nSubsets; ++subset)
    {
      const int subiter = iter * p.nSubsets + subset + 1;
      if (p.nSubsets > 1)
        print("  Sub-iteration ", subset + 1, " of ", p.nSubsets);
      const int conv =
        convolveFlag && subiter % p.convolutionInterval == 0 ? 1 : 0;
      device::check(fir_osem_subiter(
        ctx, &v, x.f(), y.f(), biasProj ? bias.f() : nullptr,
        s.f() + nv * subset, bp.f(), subset, p.nSubsets, conv,
        p.fwhmXYZ.data(), p.cutRadius, subiter != total ? 1 : 0));
      if (p.saveInterval > 0 && subiter % p.saveInterval == 0 &&
          subiter != total)
      {
        x.download(outputVol.getDataArray(), vb);
        outputVol.write(outputVolFileName + "_subiter_" +
                        std::to_string(subiter));
      }
    }
  }
  x.download(outputVol.getDataArray(), vb);
}
}
