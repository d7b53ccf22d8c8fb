// fir_host_selftest -- exercises the host layer's file formats and
// geometry without a GPU (used by the CPU test suite).
//   vol IN.h33 OUT      read a volume and write it back
//   proj IN.hs OUT      read a sinogram and write it back
//   scanner IN.scan     print the crystal/slice tables (%.17g)
//   geom R C SPAN SEG T print ProjGeometry::fill results, or the error
//   lors IN.hs NSUB     print every LOR of the cache, in order
//   bins IN.hs OUT      ProjData::getBinCoordinates of every (ring1,
//                       crystal1, ring2, crystal2), rings and crystals
//                       ascending in that nesting: int32 records (ok, seg,
//                       view, axial, tang)
//   rebin IN.hs OUT.hs  read a sinogram, ProjData::rebinWeight, write it
//                       (needs a GPU)
//   upload IN.hs REPS   read a sinogram and time REPS uploads to and
//                       downloads from the GPU of its data (one JSON line;
//                       needs a GPU)
//   key TEXT            print the standardised keyword
//   params FILE         parse an OSEM parameter file and print the values
#include <KeyParser.h>
#include <LORCache.h>
#include <ProjData.h>
#include <ScannerData.h>
#include <VolData.h>
#include <console.h>
#include <device.h>
#include <pinned.h>
#include <reconAlgos.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

int main(int argc, char** argv)
{
  try
  {
    if (argc < 2) error("mode missing");
    const std::string mode = argv[1];
    if (mode == "vol" && argc >= 4)
    {
      VolData v(argv[2], VolData::ConstructionMode::READ_DATA);
      v.write(argv[3]);
    }
    else if (mode == "proj" && argc >= 4)
    {
      ProjData p(argv[2], ProjData::ConstructionMode::READ_DATA);
      p.write(argv[3]);
    }
    else if (mode == "scanner" && argc >= 3)
    {
      ScannerData s(argv[2]);
      const auto& g = s.getGeometry();
      std::printf("%d %d\n", g.nCrystalsPerRing, g.nSlices);
      for (int i = 0; i < g.nCrystalsPerRing; ++i)
        std::printf("%.17g %.17g\n", s.getCrystalXYPositionVector()[i].x,
                    s.getCrystalXYPositionVector()[i].y);
      for (int i = 0; i < g.nSlices; ++i)
        std::printf("%.17g\n", s.getSliceZPositionVector()[i]);
    }
    else if (mode == "geom" && argc >= 7)
    {
      ProjHeader h;
      h.setDefaults();
      h.nRings = std::atoi(argv[2]);
      h.nCrystalsPerRing = std::atoi(argv[3]);
      h.segmentSpan = std::atoi(argv[4]);
      h.nSegments = std::atoi(argv[5]);
      h.nTangCoords = std::atoi(argv[6]);
      h.check();
      ProjGeometry g;
      g.fill(h);
      std::printf("%d %d %d %d %d %d %d", g.nBins, g.segOffset,
                  g.tangCoordOffset, g.nViews, g.halfSegmentSpan,
                  g.maxRingDiff, h.nTangCoords);
      for (int n : g.nAxialCoords) std::printf(" %d", n);
      std::printf("\n");
    }
    else if (mode == "bins" && argc >= 4)
    {
      ProjData p(argv[2], ProjData::ConstructionMode::ALLOCATE);
      const auto& h = p.getHeader();
      std::vector<int> out;
      for (int r1 = 0; r1 < h.nRings; ++r1)
        for (int c1 = 0; c1 < h.nCrystalsPerRing; ++c1)
          for (int r2 = 0; r2 < h.nRings; ++r2)
            for (int c2 = 0; c2 < h.nCrystalsPerRing; ++c2)
            {
              int s = 0, v = 0, a = 0, t = 0;
              const bool ok = p.getBinCoordinates(r1, c1, r2, c2, &s, &v, &a, &t);
              out.insert(out.end(), {ok ? 1 : 0, s, v, a, t});
            }
      FILE* f = std::fopen(argv[3], "wb");
      if (!f) error("cannot write ", argv[3]);
      std::fwrite(out.data(), 4, out.size(), f);
      std::fclose(f);
    }
    else if (mode == "rebin" && argc >= 4)
    {
      ProjData p(argv[2], ProjData::ConstructionMode::READ_DATA);
      p.rebinWeight();
      p.write(argv[3]);
    }
    else if (mode == "upload" && argc >= 4)
    {
      ProjData p(argv[2], ProjData::ConstructionMode::READ_DATA);
      const int reps = std::atoi(argv[3]);
      const size_t bytes =
        sizeof(float) * static_cast<size_t>(p.getGeometry().nBins);
      float* data = p.getContiguousData();
      device::Buffer b(bytes);
      b.upload(data, bytes);
      b.download(data, bytes);
      using clk = std::chrono::steady_clock;
      const auto t0 = clk::now();
      for (int r = 0; r < reps; ++r) b.upload(data, bytes);
      const auto t1 = clk::now();
      for (int r = 0; r < reps; ++r) b.download(data, bytes);
      const auto t2 = clk::now();
      const double up = std::chrono::duration<double>(t1 - t0).count() / reps;
      const double dn = std::chrono::duration<double>(t2 - t1).count() / reps;
      std::printf("{\"pinned\": %d, \"bytes\": %zu, \"h2d_gb_s\": %.2f, "
                  "\"d2h_gb_s\": %.2f}\n",
                  pinned::isPinned(data) ? 1 : 0, bytes, bytes / up / 1e9,
                  bytes / dn / 1e9);
    }
    else if (mode == "lors" && argc >= 4)
    {
      ProjData p(argv[2], ProjData::ConstructionMode::ALLOCATE);
      const int ns = std::atoi(argv[3]);
      LORCache c(p, ns);
      for (int s = 0; s < ns; ++s)
        for (int seg = -p.getGeometry().segOffset;
             seg <= p.getGeometry().segOffset; ++seg)
        {
          const int n = c.setSubsetAndSegment(s, seg);
          for (int i = 0; i < n; ++i)
          {
            const auto [ok, bin, s1, a1, s2, a2] = c.getLOR(i);
            std::printf("%d %d %d %d %d %d %d\n", seg, ok ? 1 : 0, bin, s1,
                        a1, s2, a2);
          }
        }
    }
    else if (mode == "key" && argc >= 3)
    {
      std::printf("[%s]\n", KeyParser::standardiseKeyword(argv[2]).c_str());
    }
    else if (mode == "params" && argc >= 3)
    {
      std::string in, scan, hdr, out, sens, bias, hu, acf;
      OSEMCoreParams a;
      KeyParser kp;
      kp.addStartKey("!OSEM PARAMETERS");
      kp.addKey("input projection file", &in);
      kp.addKey("scanner file", &scan);
      kp.addKey("output volume header", &hdr);
      kp.addKey("output volume file name", &out);
      kp.addKey("number of iterations", &a.nIterations);
      kp.addKey("number of subsets", &a.nSubsets);
      kp.addKey("save interval", &a.saveInterval);
      kp.addKey("cut radius in mm", &a.cutRadius);
      kp.addKey("convolution interval", &a.convolutionInterval);
      kp.addKey("convolution FHWM XYZ in mm", &a.fwhmXYZ);
      kp.addKey("sensitivity map volume", &sens);
      kp.addKey("bias projection", &bias);
      kp.addKey("attenuation volume in HU", &hu);
      kp.addKey("attenuation correction factors", &acf);
      kp.addStopKey("!END OF OSEM PARAMETERS");
      const bool ok = kp.parse(argv[2]);
      std::printf("ok=%d\nin=[%s]\nscan=[%s]\nhdr=[%s]\nout=[%s]\n", ok ? 1 : 0,
                  in.c_str(), scan.c_str(), hdr.c_str(), out.c_str());
      std::printf("iter=%d\nsub=%d\nsave=%d\ncut=%.9g\nconv=%d\nfwhm=",
                  a.nIterations, a.nSubsets, a.saveInterval, a.cutRadius,
                  a.convolutionInterval);
      for (float f : a.fwhmXYZ) std::printf("%.9g,", f);
      std::printf("\nsens=[%s]\nbias=[%s]\nhu=[%s]\nacf=[%s]\n", sens.c_str(),
                  bias.c_str(), hu.c_str(), acf.c_str());
    }
    else
      error("bad arguments");
  }
  catch (const std::exception& ex)
  {
    std::cerr << ex.what();
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
