// ref_driver.cc -- small harness around the compiled reference library.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile against the reference
// sources under /root/reference (never copied into this repository) into
// oracle/_ref/. Used to produce golden vectors and, in bench.py's
// cpu_baseline leg, to time the reference's own OpenMP operators in-process
// (file I/O excluded).
//
// Modes:
//   segpath nx ny nz dx dy dz ox oy oz x1 y1 z1 x2 y2 z2
//       print the Siddon path of one segment ("n idx len idx len ...",
//       n = -1 for a miss) with lengths in %.17g.
//   paths vol.h33 scanner.scan proj.hs out.bin [b0 nb]
//       binary dump of the paths of bins [b0, b0+nb) in segment-concatenated
//       bin order: int64 nb, int64 total, int32 counts[nb],
//       int32 idx[total], float64 len[total].
//   tables scanner.scan out.bin
//       int32 nCrystalsPerRing, int32 nSlices, float64 xy[2n], float64 z[s].
//   bins proj.hs out.bin
//       ProjData::getBinCoordinates of every (ring1, crystal1, ring2,
//       crystal2), rings and crystals ascending in that nesting: int32
//       records (ok, seg, view, axial, tang).
//   rebin proj.hs out.bin
//       ProjData::rebinWeight of a sinogram of ones (single-segment
//       headers only: the reference's bin addressing offsets the segment
//       twice, ProjData.inl:67-82); float32 bins in file order.
//   reso vol.h33 scanner.scan proj.hs nIter nSub fwhm conv cut out.bin
//       reconAlgos::OSEM_ResoReco (no executable reaches it: OSEM.cc:341)
//       from the image of vol.h33, with the sensitivity computed as
//       FIR_OSEM computes it; writes the float32 image.
//   time op vol.h33 scanner.scan proj.hs nSubsets reps
//       op in {fwd, bp, sens, subiter}; uniform image and sinogram of ones;
//       prints one JSON line with the median seconds over reps runs after a
//       warm-up, and the thread count.
#include <LORCache.h>
#include <ProjData.h>
#include <ScannerData.h>
#include <Siddon.h>
#include <VolData.h>
#include <macros.h>
#include <projections.h>
#include <reconAlgos.h>
#include <tools.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace
{
int segpath(char** a)
{
  VolHeader h;
  h.setDefaults();
  h.volSize = {std::atoi(a[0]), std::atoi(a[1]), std::atoi(a[2])};
  h.voxelExtent = {std::atof(a[3]), std::atof(a[4]), std::atof(a[5])};
  h.volOffset = {std::atof(a[6]), std::atof(a[7]), std::atof(a[8])};
  h.nFrames = 1;
  VolData vol(h);
  Siddon siddon(vol);
  auto* pe = siddon.getThreadLocalPathElements();
  const bool hit = siddon.computePath(
    std::atof(a[9]), std::atof(a[10]), std::atof(a[11]), std::atof(a[12]),
    std::atof(a[13]), std::atof(a[14]), pe);
  if (!hit)
  {
    std::printf("-1\n");
    return 0;
  }
  int n = 0;
  while (pe[n].coord != -1) ++n;
  std::printf("%d", n);
  for (int i = 0; i < n; ++i)
    std::printf(" %d %.17g", pe[i].coord, pe[i].length);
  std::printf("\n");
  return 0;
}

int paths(int argc, char** a)
{
  VolData vol(a[0], VolData::ConstructionMode::ALLOCATE);
  ScannerData scanner(a[1]);
  ProjData proj(a[2], ProjData::ConstructionMode::ALLOCATE);
  const auto& g = proj.getGeometry();
  const auto& ph = proj.getHeader();
  long long nbins = g.nBins;
  long long b0 = 0, nb = nbins;
  if (argc >= 6)
  {
    b0 = std::atoll(a[4]);
    nb = std::atoll(a[5]);
  }
  Siddon siddon(vol);
  auto* pe = siddon.getThreadLocalPathElements();
  std::vector<int> counts;
  std::vector<int> idx;
  std::vector<double> len;
  long long b = 0;
  LOOP_SEG(seg, proj)
  {
    LOOP_VIEW(view, proj)
    LOOP_AXIAL(axial, proj, seg)
    LOOP_TANG(tang, proj)
    {
      if (b >= b0 && b < b0 + nb)
      {
        const auto [s1, s2] = proj.getCrystalAxialCoord(seg, axial);
        const auto [c1, c2] = proj.getCrystalAngCoord(view, tang);
        const bool hit =
          siddon.computePathBetweenCrystals(scanner, s1, c1, s2, c2, pe);
        int n = 0;
        while (pe[n].coord != -1)
        {
          idx.push_back(pe[n].coord);
          len.push_back(pe[n].length);
          ++n;
        }
        counts.push_back(hit ? n : -1);
      }
      ++b;
    }
  }
  (void)ph;
  FILE* f = std::fopen(a[3], "wb");
  if (!f) return 2;
  long long tot = (long long)idx.size();
  long long ncount = (long long)counts.size();
  std::fwrite(&ncount, 8, 1, f);
  std::fwrite(&tot, 8, 1, f);
  std::fwrite(counts.data(), 4, counts.size(), f);
  std::fwrite(idx.data(), 4, idx.size(), f);
  std::fwrite(len.data(), 8, len.size(), f);
  std::fclose(f);
  return 0;
}

int bins(char** a)
{
  ProjData proj(a[0], ProjData::ConstructionMode::ALLOCATE);
  const auto& ph = proj.getHeader();
  std::vector<int> out;
  for (int r1 = 0; r1 < ph.nRings; ++r1)
    for (int c1 = 0; c1 < ph.nCrystalsPerRing; ++c1)
      for (int r2 = 0; r2 < ph.nRings; ++r2)
        for (int c2 = 0; c2 < ph.nCrystalsPerRing; ++c2)
        {
          int seg = 0, view = 0, ax = 0, t = 0;
          const bool ok =
            proj.getBinCoordinates(r1, c1, r2, c2, &seg, &view, &ax, &t);
          out.insert(out.end(), {ok ? 1 : 0, seg, view, ax, t});
        }
  FILE* f = std::fopen(a[1], "wb");
  if (!f) return 2;
  std::fwrite(out.data(), 4, out.size(), f);
  std::fclose(f);
  return 0;
}

int rebin(char** a)
{
  ProjData proj(a[0], ProjData::ConstructionMode::INITIALIZE, 1.0f);
  const auto& g = proj.getGeometry();
  if (g.segOffset != 0)
  {
    std::fprintf(stderr, "rebin: single-segment headers only\n");
    return 3;
  }
  proj.rebinWeight();
  FILE* f = std::fopen(a[1], "wb");
  if (!f) return 2;
  std::fwrite(proj.getDataArray()[0], 4, g.nBins, f);
  std::fclose(f);
  return 0;
}

int reso(char** a)
{
  VolData vol(a[0], VolData::ConstructionMode::READ_DATA);
  ScannerData scanner(a[1]);
  ProjData proj(a[2], ProjData::ConstructionMode::READ_DATA);
  OSEMCoreParams p;
  p.nIterations = std::atoi(a[3]);
  p.nSubsets = std::atoi(a[4]);
  const float fw = (float)std::atof(a[5]);
  p.fwhmXYZ = {fw, fw, fw};
  p.convolutionInterval = std::atoi(a[6]);
  p.cutRadius = (float)std::atof(a[7]);
  VolData sens;
  sens.allocateAsMultiVol(vol, p.nSubsets);
  projections::computeSensitivityVol(proj, scanner, sens, p.nSubsets);
  reconAlgos::OSEM_ResoReco(proj, scanner, vol, "unused", p, sens,
                            std::nullopt);
  FILE* f = std::fopen(a[8], "wb");
  if (!f) return 2;
  std::fwrite(vol.getDataArray(), 4, (size_t)vol.getNVoxelsPerFrame(), f);
  std::fclose(f);
  return 0;
}

int tables(char** a)
{
  ScannerData scanner(a[0]);
  // The slice count is 2*nRings-1; ring and crystal counts come from a
  // probe ProjData-free route: scan the xy table length via the header.
  const int n = std::atoi(a[2]);
  const int ns = std::atoi(a[3]);
  FILE* f = std::fopen(a[1], "wb");
  if (!f) return 2;
  std::fwrite(&n, 4, 1, f);
  std::fwrite(&ns, 4, 1, f);
  const auto* xy = scanner.getCrystalXYPositionVector();
  for (int i = 0; i < n; ++i)
  {
    std::fwrite(&xy[i].x, 8, 1, f);
    std::fwrite(&xy[i].y, 8, 1, f);
  }
  std::fwrite(scanner.getSliceZPositionVector(), 8, ns, f);
  std::fclose(f);
  return 0;
}

double now()
{
  return std::chrono::duration<double>(
           std::chrono::steady_clock::now().time_since_epoch())
    .count();
}

int timeit(char** a)
{
  const std::string op = a[0];
  VolData vol(a[1], VolData::ConstructionMode::INITIALIZE, 1.0);
  ScannerData scanner(a[2]);
  ProjData proj(a[3], ProjData::ConstructionMode::INITIALIZE, 1.0);
  const int nSubsets = std::atoi(a[4]);
  const int reps = std::atoi(a[5]);
  std::vector<double> t;
  VolData frames;
  if (op == "bp" || op == "sens")
    frames.allocateAsMultiVol(vol, op == "bp" ? 1 : nSubsets);
  VolData sens;
  OSEMCoreParams params;
  params.nIterations = 1;
  params.nSubsets = nSubsets;
  if (op == "subiter")
  {
    // One sub-iteration = OSEM with a single iteration over one subset of
    // nSubsets; the reference loops all subsets inside OSEM(), so time one
    // full iteration and report per-sub-iteration time as well.
    sens.allocateAsMultiVol(vol, nSubsets);
    sens.setAllVoxelsAllFrames(1.0);
  }
  for (int r = 0; r <= reps; ++r)
  {
    const double t0 = now();
    if (op == "fwd")
      projections::forward(vol, scanner, proj);
    else if (op == "bp")
      projections::backward(proj, scanner, frames, 1);
    else if (op == "sens")
      projections::computeSensitivityVol(proj, scanner, frames, nSubsets);
    else if (op == "subiter")
    {
      vol.setAllVoxels(1.0);
      reconAlgos::OSEM(proj, scanner, vol, "/dev/null", params, sens,
                       std::nullopt);
    }
    else
      return 3;
    const double t1 = now();
    if (r > 0) t.push_back(t1 - t0);
  }
  std::sort(t.begin(), t.end());
  std::printf(
    "{\"op\": \"%s\", \"median_s\": %.6f, \"min_s\": %.6f, \"reps\": %d, "
    "\"threads\": %d, \"n_subsets\": %d, \"n_bins\": %d}\n",
    op.c_str(), t[t.size() / 2], t[0], reps, getNThreads(), nSubsets,
    proj.getGeometry().nBins);
  return 0;
}
}

int main(int argc, char** argv)
{
  try
  {
    if (argc < 2) return 1;
    const std::string mode = argv[1];
    if (mode == "segpath" && argc >= 17) return segpath(argv + 2);
    if (mode == "paths" && argc >= 6) return paths(argc - 2, argv + 2);
    if (mode == "tables" && argc >= 6) return tables(argv + 2);
    if (mode == "bins" && argc >= 4) return bins(argv + 2);
    if (mode == "rebin" && argc >= 4) return rebin(argv + 2);
    if (mode == "reso" && argc >= 11) return reso(argv + 2);
    if (mode == "time" && argc >= 8) return timeit(argv + 2);
    std::fprintf(stderr, "bad arguments\n");
    return 1;
  }
  catch (const std::exception& ex)
  {
    std::fprintf(stderr, "%s", ex.what());
    return 1;
  }
}
