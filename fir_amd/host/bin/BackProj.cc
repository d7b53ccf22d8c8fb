// FIR_BackProj proj.hs scanner.scan volTemplate.h33 outName -- single-frame
// backprojection (reference src_bin/BackProj.cc:8-48), on the GPU.
// Requires all four arguments (the reference checks for three but reads
// the fourth).
#include <ProjData.h>
#include <ScannerData.h>
#include <VolData.h>
#include <console.h>
#include <projections.h>

#include <cstdlib>
#include <iostream>

int main(int argc, char** argv)
{
  try
  {
    if (argc < 5) error("Requires four input arguments");
    ProjData inputProj(argv[1], ProjData::ConstructionMode::READ_DATA);
    ScannerData scanner(argv[2]);
    VolData outputVol(argv[3], VolData::ConstructionMode::ALLOCATE);
    projections::backward(inputProj, scanner, outputVol);
    outputVol.write(argv[4]);
  }
  catch (const std::exception& ex)
  {
    std::cerr << ex.what();
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
