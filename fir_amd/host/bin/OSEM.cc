// FIR_OSEM params [recomputeSens] [recomputeAttn] -- OSEM reconstruction
// with the reference's parameter file, flags, sensitivity/ACF caching and
// outputs (src_bin/OSEM.cc:16-502); all projector work runs on the GPU.
//
// Flags: anything not starting with '0' means 1; both default to 1.
//   recomputeSens: recompute the sensitivity map (always recomputed when no
//     sensitivity file is named or the named file does not exist; saved when
//     a file is named).
//   recomputeAttn: recompute the attenuation correction factors from the HU
//     volume (falls back to reading the ACF file when no HU volume is
//     given, and to recomputing when the ACF file does not exist).
#include <KeyParser.h>
#include <ProjData.h>
#include <ScannerData.h>
#include <VolData.h>
#include <console.h>
#include <operations.h>
#include <projections.h>
#include <reconAlgos.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <optional>
#include <string>

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
    ProjData inputProj(params.inputProjFile,
                       ProjData::ConstructionMode::READ_DATA);
    inputProj.checkNSubsets(params.algo.nSubsets);
    ScannerData scanner(params.scannerFile);

    // initial estimate: the template's data, or ones
    VolData outputVol(params.outputVolHeader,
                      VolData::ConstructionMode::READ_DATA_IF_PROVIDED, 1.0);
    printEmptyLine();
    tRead.stop();

    VolData sensVol;
    PhaseTimer tSens("sensitivity");
    if (recomputeSens)
    {
      echo("Computing sensitivity map");
      printEmptyLine();
      sensVol.allocateAsMultiVol(outputVol, params.algo.nSubsets);
      projections::computeSensitivityVol(inputProj, scanner, sensVol,
                                         params.algo.nSubsets);
      if (sensNamed)
      {
        printQuotedValue("Saving sensitivity map to file", params.sensVolFile);
        printEmptyLine();
        sensVol.write(params.sensVolFile);
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
    }

    tSens.stop();

    std::optional<ProjData> biasProj;
    if (!params.biasProjFile.empty())
    {
      printQuotedValue("Reading bias projection from file",
                       params.biasProjFile);
      printEmptyLine();
      biasProj.emplace(params.biasProjFile,
                       ProjData::ConstructionMode::READ_DATA);
    }

    PhaseTimer tAcf("attenuation");
    if (huNamed || acfExists)
    {
      if (!recomputeAcf && huNamed && !acfExists) recomputeAcf = true;
      if (recomputeAcf && !huNamed && acfExists) recomputeAcf = false;
      ProjData acf;
      if (recomputeAcf)
      {
        echo("Computing attenuation correction factors");
        printEmptyLine();
        VolData muMap(params.attenVolHUFile,
                      VolData::ConstructionMode::READ_DATA);
        operations::HounsfieldToMuMap(muMap);
        operations::cutCircle(muMap, params.algo.cutRadius);
        acf.copy(inputProj, ProjData::ConstructionMode::INITIALIZE);
        projections::forward(muMap, scanner, acf);
        acf.exponential();
        if (acfNamed)
        {
          printQuotedValue("Saving attenuation correction factors to file",
                           params.attenCorrFactorsFile);
          printEmptyLine();
          acf.write(params.attenCorrFactorsFile);
        }
      }
      else
      {
        printQuotedValue("Reading attenuation correction factors from file",
                         params.attenCorrFactorsFile);
        printEmptyLine();
        acf.read(params.attenCorrFactorsFile,
                 ProjData::ConstructionMode::READ_DATA);
      }
      inputProj *= acf;
    }

    tAcf.stop();

    PhaseTimer tOsem("OSEM");
    // The reference selects OSEM_ResoReco with a constant that is always
    // false (OSEM.cc:340-341); here FIR_RESO_RECO=1 selects it.
    const char* reso = std::getenv("FIR_RESO_RECO");
    if (reso && *reso && std::strcmp(reso, "0") != 0)
      reconAlgos::OSEM_ResoReco(inputProj, scanner, outputVol,
                                params.outputVolFileName, params.algo, sensVol,
                                biasProj);
    else
      reconAlgos::OSEM(inputProj, scanner, outputVol,
                       params.outputVolFileName, params.algo, sensVol,
                       biasProj);

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
