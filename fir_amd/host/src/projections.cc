#include <console.h>
#include <device.h>
#include <projections.h>

namespace projections
{
void forward(const VolData& in, const ScannerData& scanner, ProjData& out)
{
  scanner.checkProjData(out);
  device::useGeometry(scanner, out.getHeader());
  const fir_vol_desc v = device::describe(in.getHeader());
  const size_t nv = static_cast<size_t>(in.getNVoxelsPerFrame());
  const size_t nb = static_cast<size_t>(out.getGeometry().nBins);
  device::Buffer img(nv * sizeof(float)), sino(nb * sizeof(float));
  img.upload(in.getDataArray(), nv * sizeof(float));
  echo("Forward projection (GPU)");
  device::check(fir_forward(device::context(), &v, img.f(), sino.f()));
  sino.download(out.getContiguousData(), nb * sizeof(float));
}

void backward(const ProjData& in, const ScannerData& scanner, VolData& out,
              int nSubsets)
{
  scanner.checkProjData(in);
  in.checkNSubsets(nSubsets);
  out.checkNFrames(nSubsets);
  device::useGeometry(scanner, in.getHeader());
  const fir_vol_desc v = device::describe(out.getHeader());
  const size_t nv = static_cast<size_t>(out.getNVoxelsPerFrame());
  const size_t nb = static_cast<size_t>(in.getGeometry().nBins);
  device::Buffer sino(nb * sizeof(float));
  device::Buffer img(nv * nSubsets * sizeof(float));
  sino.upload(in.getContiguousData(), nb * sizeof(float));
  echo("Back-projection (GPU)");
  device::check(
    fir_backward(device::context(), &v, sino.f(), img.f(), nSubsets));
  img.download(out.getAllFrames(), nv * nSubsets * sizeof(float));
}

void computeSensitivityVol(const ProjData& proj, const ScannerData& scanner,
                           VolData& sens, int nSubsets)
{
  scanner.checkProjData(proj);
  proj.checkNSubsets(nSubsets);
  sens.checkNFrames(nSubsets);
  device::useGeometry(scanner, proj.getHeader());
  const fir_vol_desc v = device::describe(sens.getHeader());
  const size_t nv = static_cast<size_t>(sens.getNVoxelsPerFrame());
  device::Buffer img(nv * nSubsets * sizeof(float));
  echo("Sensitivity (GPU)");
  device::check(
    fir_sensitivity(device::context(), &v, img.f(), nSubsets));
  img.download(sens.getAllFrames(), nv * nSubsets * sizeof(float));
}
}
