// FIR_ForwardProj params -- forward projection of a volume's frames.
// Same parameter file and outputs as the reference executable
// (src_bin/ForwardProj.cc:16-197); the projection runs on the GPU.
#include <KeyParser.h>
#include <ProjData.h>
#include <ScannerData.h>
#include <VolData.h>
#include <console.h>
#include <device.h>
#include <operations.h>

#include <cstdlib>
#include <iostream>
#include <numeric>
#include <optional>
#include <string>
#include <vector>

namespace
{
struct Params
{
  explicit Params(const char* file);
  // Fill `frames` with every frame if empty; returns true when a single
  // frame of a single-frame volume is projected (output name unsuffixed).
  bool resolveFrames(int nFrames);

  std::string inputVolFile, scannerFile, outputProjHeader, outputProjFileName;
  std::vector<int> frames;
  std::string maskVolFile;
};

Params::Params(const char* file)
{
  KeyParser kp;
  kp.addStartKey("!FORWARD PROJECTION PARAMETERS");
  kp.addKey("input volume file", &inputVolFile);
  kp.addKey("scanner file", &scannerFile);
  kp.addKey("output projection header", &outputProjHeader);
  kp.addKey("output projection file name", &outputProjFileName);
  kp.addKey("frames to project", &frames);
  kp.addKey("mask volume file", &maskVolFile);
  kp.addStopKey("!END OF FORWARD PROJECTION PARAMETERS");
  kp.parse(file);
  if (inputVolFile.empty()) error("No input volume file provided");
  if (scannerFile.empty()) error("No scanner file provided");
  if (outputProjHeader.empty()) error("No output projection header provided");
  if (outputProjFileName.empty())
    error("No output projection file name provided");
}

bool Params::resolveFrames(int nFrames)
{
  if (frames.empty())
  {
    frames.resize(nFrames);
    std::iota(frames.begin(), frames.end(), 0);
    return nFrames == 1;
  }
  for (int f : frames)
    if (f < 0 || f >= nFrames)
      error("Frame indices must be between 0 and ", nFrames - 1);
  return false;
}
}

int main(int argc, char** argv)
{
  try
  {
    printEmptyLine();
    echo("=== FIR_ForwardProj (MI355X) ===");
    printEmptyLine();
    if (argc < 2) error("Parameter file missing");
    Params params(argv[1]);

    VolData inputVol(params.inputVolFile, VolData::ConstructionMode::READ_DATA);
    const bool single = params.resolveFrames(inputVol.getNFrames());
    std::optional<VolData> mask;
    if (!params.maskVolFile.empty())
      mask.emplace(params.maskVolFile, VolData::ConstructionMode::READ_DATA);
    ScannerData scanner(params.scannerFile);
    ProjData out(params.outputProjHeader, ProjData::ConstructionMode::ALLOCATE);

    // projections::forward (projections.cc:15-92) on the device; the
    // sinogram is downloaded chunk by chunk while earlier chunks are written
    scanner.checkProjData(out);
    device::useGeometry(scanner, out.getHeader());
    const fir_vol_desc v = device::describe(inputVol.getHeader());
    const size_t nv = static_cast<size_t>(inputVol.getNVoxelsPerFrame());
    const size_t nb = static_cast<size_t>(out.getGeometry().nBins);
    device::Buffer img(nv * sizeof(float)), sino(nb * sizeof(float));
    for (int frame : params.frames)
    {
      inputVol.setActiveFrame(frame);
      if (mask) operations::applyMask(inputVol, *mask);
      img.upload(inputVol.getDataArray(), nv * sizeof(float));
      echo("Forward projection (GPU)");
      device::check(fir_forward(device::context(), &v, img.f(), sino.f()));
      device::writeProj(sino, out,
                        single ? params.outputProjFileName
                               : params.outputProjFileName + "_frame_" +
                                   std::to_string(frame));
    }
  }
  catch (const std::exception& ex)
  {
    std::cerr << ex.what();
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
