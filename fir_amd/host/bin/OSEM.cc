// FIR_OSEM params [recomputeSens] [recomputeAttn] -- OSEM reconstruction
// with the reference's parameter file, flags, sensitivity/ACF caching and
// outputs (src_bin/OSEM.cc:16-502); all projector work runs on the GPU,
// and the sinograms stay in device memory from the moment they are read
// (each chunk's upload overlaps the reading of the next).
//
// Flags: anything not starting with '0' means 1; both default to 1.
//   recomputeSens: recompute the sensitivity map (always recomputed when no
//     sensitivity file is named or the named file does not exist; saved when
//     a file is named). It is the sensitivity of the projector the
//     iterations use: with FIR_PATH_CACHE, that of the cached paths.
//   recomputeAttn: recompute the attenuation correction factors from the HU
//     volume (falls back to reading the ACF file when no HU volume is
//     given, and to recomputing when the ACF file does not exist).
// With FIR_PATH_CACHE the path cache is built on a second host thread
// while the sinogram and maps are read (it depends on the geometry only).
#include <KeyParser.h>
#include <ProjData.h>
#include <ProjInterfileReader.h>
#include <ScannerData.h>
#include <VolData.h>
#include <console.h>
#include <device.h>
#include <reconAlgos.h>

#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <thread>

namespace
{
struct Params
{
  explicit Params(const char* file);
  std::string inputProjFile, scannerFile, outputVolHeader, outputVolFileName;
  OSEMCoreParams algo;
  std::string sensVolFile, biasProjFile, attenVolHUFile, attenCorrFactorsFile;
};

Params::Params(const char* file)
{
  KeyParser kp;
  kp.addStartKey("!OSEM PARAMETERS");
  kp.addKey("input projection file", &inputProjFile);
  kp.addKey("scanner file", &scannerFile);
  kp.addKey("output volume header", &outputVolHeader);
  kp.addKey("output volume file name", &outputVolFileName);
  kp.addKey("number of iterations", &algo.nIterations);
  kp.addKey("number of subsets", &algo.nSubsets);
  kp.addKey("save interval", &algo.saveInterval);
  kp.addKey("cut radius in mm", &algo.cutRadius);
  kp.addKey("convolution interval", &algo.convolutionInterval);
  kp.addKey("convolution FHWM XYZ in mm", &algo.fwhmXYZ);
  kp.addKey("sensitivity map volume", &sensVolFile);
  kp.addKey("bias projection", &biasProjFile);
  kp.addKey("attenuation volume in HU", &attenVolHUFile);
  kp.addKey("attenuation correction factors", &attenCorrFactorsFile);
  kp.addStopKey("!END OF OSEM PARAMETERS");
  kp.parse(file);
  if (inputProjFile.empty()) error("No input projection file provided");
  if (scannerFile.empty()) error("No scanner file provided");
  if (outputVolHeader.empty()) error("No output volume header provided");
  if (outputVolFileName.empty()) error("No output volume file name provided");
}

bool fileExists(const std::string& path)
{
  std::ifstream f(path);
  return f.good();
}

bool flag(int argc, char** argv, int i)
{
  return argc > i ? argv[i][0] != '0' : true;
}
}

// The path cache built on a second host thread (reconAlgos::buildPathCache
// on the shared device context, whose stream orders its work with the
// reads' uploads); the thread is joined on every exit path and its error,
// if any, is raised by finish().
class CacheBuilder
{
public:
  void start(const fir_vol_desc& v, int nSubsets)
  {
    mThread = std::thread([this, v, nSubsets] {
      try
      {
        mCache = reconAlgos::buildPathCache(v, nSubsets);
      }
      catch (...)
      {
        mError = std::current_exception();
      }
    });
  }
  fir_cache* finish()
  {
    if (mThread.joinable()) mThread.join();
    if (mError) std::rethrow_exception(mError);
    return mCache ? mCache->cache : nullptr;
  }
  ~CacheBuilder()
  {
    if (mThread.joinable()) mThread.join();
  }

private:
  std::thread mThread;
  std::unique_ptr<reconAlgos::PathCacheHandle> mCache;
  std::exception_ptr mError;
};

int main(int argc, char** argv)
{
  try
  {
    printEmptyLine();
    echo("=== FIR_OSEM (MI355X) ===");
    printEmptyLine();
    if (argc < 2) error("Parameter file missing");
    Params params(argv[1]);

    const bool sensNamed = !params.sensVolFile.empty();
    const bool huNamed = !params.attenVolHUFile.empty();
    const bool acfNamed = !params.attenCorrFactorsFile.empty();
    bool recomputeSens = flag(argc, argv, 2);
    bool recomputeAcf = flag(argc, argv, 3);
    if (!recomputeSens)
      recomputeSens = !sensNamed || !fileExists(params.sensVolFile);
    const bool acfExists = acfNamed && fileExists(params.attenCorrFactorsFile);

    PhaseTimer total("total");
    PhaseTimer tRead("read inputs");
    // geometry first (scanner, sinogram header, image grid): a path cache
    // depends on nothing else, so it is built on a second host thread
    // while the data are read
    ScannerData scanner(params.scannerFile);
    const ProjHeader projHeader =
      ProjInterfileReader(params.inputProjFile).getHeader();
    // initial estimate: the template's data, or ones
    VolData outputVol(params.outputVolHeader,
                      VolData::ConstructionMode::READ_DATA_IF_PROVIDED, 1.0);
    device::useGeometry(scanner, projHeader);
    fir_ctx* ctx = device::context();
    const int nSub = params.algo.nSubsets;
    CacheBuilder builder;
    if (reconAlgos::pathCacheRequested())
    {
      const fir_vol_desc v = device::describe(outputVol.getHeader());
      builder.start(v, nSub);
    }

    // the sinogram goes to the device chunk by chunk as it is read; every
    // step below keeps its data there
    ProjData inputProj;
    device::Buffer y = device::readProj(params.inputProjFile, inputProj);
    inputProj.checkNSubsets(nSub);
    scanner.checkProjData(inputProj);
    printEmptyLine();
    tRead.stop();
    const size_t vb =
      static_cast<size_t>(outputVol.getNVoxelsPerFrame()) * sizeof(float);
    const size_t nb = static_cast<size_t>(inputProj.getGeometry().nBins);

    device::Buffer sens(vb * nSub);
    VolData sensVol;
    DeviceOSEMInputs in;
    in.y = y.f();
    in.sens = sens.f();
    if (recomputeSens)
    {
      // computed on the device with the projector the iterations use,
      // right before them (after the path cache, when one is built)
      echo("Computing sensitivity map");
      printEmptyLine();
      in.computeSens = true;
      if (sensNamed)
      {
        sensVol.allocateAsMultiVol(outputVol, nSub);
        in.sensOut = &sensVol;
      }
    }
    else
    {
      printQuotedValue("Reading sensitivity map from file",
                       params.sensVolFile);
      printEmptyLine();
      sensVol.read(params.sensVolFile, VolData::ConstructionMode::READ_DATA);
      if (sensVol.getHeader() != outputVol.getHeader())
        error("Sensitivity volume provided doesn't fit with output volume "
              "provided");
      if (sensVol.getNFrames() != nSub)
        error("Wrong number of frames (expected ", nSub, ")");
      sens.upload(sensVol.getAllFrames(), vb * nSub);
    }

    ProjData biasProj;
    device::Buffer bias;
    if (!params.biasProjFile.empty())
    {
      printQuotedValue("Reading bias projection from file",
                       params.biasProjFile);
      printEmptyLine();
      bias = device::readProj(params.biasProjFile, biasProj);
      if (biasProj.getHeader() != inputProj.getHeader())
        error("Bias projection doesn't fit with input projection");
      in.bias = bias.f();
    }

    // the cache build is done before the context is used for projections
    in.cache = builder.finish();

    PhaseTimer tAcf("attenuation");
    if (huNamed || acfExists)
    {
      if (!recomputeAcf && huNamed && !acfExists) recomputeAcf = true;
      if (recomputeAcf && !huNamed && acfExists) recomputeAcf = false;
      ProjData acf;
      device::Buffer acfDev;
      if (recomputeAcf)
      {
        // the reference's chain (OSEM.cc:294-337): HU -> mu, cut, forward
        // projection, clamped exponential; all on the device
        echo("Computing attenuation correction factors");
        printEmptyLine();
        VolData muMap(params.attenVolHUFile,
                      VolData::ConstructionMode::READ_DATA);
        const fir_vol_desc mv = device::describe(muMap.getHeader());
        const size_t mn = static_cast<size_t>(muMap.getNVoxelsPerFrame());
        device::Buffer mu(mn * sizeof(float));
        mu.upload(muMap.getDataArray(), mn * sizeof(float));
        device::check(fir_hu_to_mu(ctx, mu.f(), static_cast<int64_t>(mn)));
        device::check(
          fir_cut_circle(ctx, &mv, mu.f(), 1, params.algo.cutRadius));
        acfDev = device::Buffer(nb * sizeof(float));
        echo("Forward projection (GPU)");
        device::check(fir_forward(ctx, &mv, mu.f(), acfDev.f()));
        device::check(
          fir_exp_clamped(ctx, acfDev.f(), static_cast<int64_t>(nb)));
        if (acfNamed)
        {
          printQuotedValue("Saving attenuation correction factors to file",
                           params.attenCorrFactorsFile);
          printEmptyLine();
          acf.copy(inputProj, ProjData::ConstructionMode::ALLOCATE);
          device::writeProj(acfDev, acf, params.attenCorrFactorsFile);
        }
      }
      else
      {
        printQuotedValue("Reading attenuation correction factors from file",
                         params.attenCorrFactorsFile);
        printEmptyLine();
        acfDev = device::readProj(params.attenCorrFactorsFile, acf);
        if (acf.getHeader() != inputProj.getHeader())
          error("Projections do not have the same size");
      }
      // inputProj *= ACF (ProjData.cc:481-512), on the device copy
      device::check(fir_mul_clamped(ctx, y.f(), acfDev.f(),
                                    static_cast<int64_t>(nb)));
      device::check(fir_synchronize(ctx));
    }
    tAcf.stop();

    PhaseTimer tOsem("OSEM");
    // The reference selects OSEM_ResoReco with a constant that is always
    // false (OSEM.cc:340-341); here FIR_RESO_RECO=1 selects it.
    const char* reso = std::getenv("FIR_RESO_RECO");
    const bool resoReco = reso && *reso && std::strcmp(reso, "0") != 0;
    reconAlgos::OSEMOnDevice(inputProj, scanner, outputVol,
                             params.outputVolFileName, params.algo, in,
                             resoReco);
    if (in.sensOut)
    {
      printQuotedValue("Saving sensitivity map to file", params.sensVolFile);
      printEmptyLine();
      sensVol.write(params.sensVolFile);
    }
    tOsem.stop();

    PhaseTimer tWrite("write output");
    printQuotedValue("Saving reconstructed volume to file",
                     params.outputVolFileName);
    printEmptyLine();
    outputVol.write(params.outputVolFileName);
  }
  catch (const std::exception& ex)
  {
    std::cerr << ex.what();
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
