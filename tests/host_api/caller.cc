// A caller written against the reference FIR library's public C++ API only
// (ScannerData.h, VolData.h, types.h): compiled against the reference
// (tests/golden/make_golden.py hostapi, which records its output) and
// against fir_amd/host/include + libfir_host.so (tests/test_host_api.py,
// which compares the output with that record). It runs on the host only.
//
//   caller SCANNER.scan VOLUME.h33 N_RSECTORS N_MODULES N_CRYSTALS
#include <ScannerData.h>
#include <VolData.h>
#include <types.h>

#include <cstdio>
#include <cstdlib>
#include <iostream>

int main(int argc, char** argv)
{
  if (argc < 6) return 1;
  try
  {
    ScannerData scanner(argv[1]);
    VolData vol(argv[2]);
    const int nr = std::atoi(argv[3]), nm = std::atoi(argv[4]),
              nc = std::atoi(argv[5]);
    for (int r = 0; r < nr; ++r)
      for (int m = 0; m < nm; ++m)
        for (int c = 0; c < nc; ++c)
        {
          int ring = -1, crystal = -1;
          scanner.getCrystalCoordinates(r, m, c, &ring, &crystal);
          std::printf("crystal %d %d %d -> ring %d crystal %d\n", r, m, c,
                      ring, crystal);
        }
    types::PathElement path[] = {{2717, 1.5},
                                 {23395, 0.25},
                                 {4000, 2.0},
                                 {23395, 0.7071067811865476},
                                 {31849, 1e-3},
                                 {-1, 0.0}};
    const types::VoxelValue line = vol.computeLineIntegral(path);
    std::printf("line %.9g\n", line);
    vol.projectLineIntegral(path, 2.5f);
    const int probe[] = {2717, 23395, 4000, 31849, 5};
    for (int v : probe)
      std::printf("voxel %d %.9g\n", v, vol.getDataArray()[v]);
    std::fflush(stdout);
    scanner.printBinInfo(1, 2, 3, -4, 5, 1, 40, 3, vol, path, line);
    scanner.printBinInfo(0, 0, 0, 0, 0, 0, 1, 0, vol);
    std::cout << std::flush;
  }
  catch (const std::exception& e)
  {
    std::fprintf(stderr, "%s", e.what());
    return 2;
  }
  return 0;
}
