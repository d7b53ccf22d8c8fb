#include <ProjData.h>
#include <ScannerData.h>
#include <ScannerInterfileReader.h>
#include <VolData.h>
#include <console.h>
#include <macros.h>

#include <cmath>
#include <cstdio>
#include <iostream>

ScannerData::ScannerData(const std::string& file)
{
  readScannerInterfileHeader(file, mHeader);
  mGeometry.fill(mHeader);
  computeTables();
}

ScannerData::ScannerData(const ScannerHeader& header) : mHeader(header)
{
  mHeader.check();
  mGeometry.fill(mHeader);
  computeTables();
}

void ScannerData::checkProjData(const ProjData& proj) const
{
  const auto& h = proj.getHeader();
  if (h.nRings != mGeometry.nRings ||
      h.nCrystalsPerRing != mGeometry.nCrystalsPerRing)
    error("Projection must have the same number of rings (", mGeometry.nRings,
          ") and the same number of crystals per ring (",
          mGeometry.nCrystalsPerRing, ") as the scanner");
}

// Positions as the reference builds them (ScannerData.cc:234-377): the
// first rSector is vertical at x = inner radius + depth/2, its crystals
// centred on y = 0; crystal 0 is the middle crystal of that rSector and
// the others follow by rotating whole rSectors by 2*pi/rSectorRepeat.
void ScannerData::computeTables()
{
  const auto& h = mHeader;
  const auto& g = mGeometry;
  const int ny = h.crystalRepeatNumbersYZ[0];
  const int my = h.moduleRepeatNumbersYZ[0];
  const double cvy = g.crystalRepeatVectorYZ[0];
  const double mvy = g.moduleRepeatVectorYZ[0];
  const double y0 = -((my - 1) * mvy + (ny - 1) * cvy) / 2.0;
  const int perSector = my * ny;
  std::vector<double> sectorY(perSector);
  for (int m = 0; m < my; ++m)
    for (int c = 0; c < ny; ++c)
      sectorY[m * ny + c] = m * mvy + c * cvy + y0;

  mXY.assign(g.nCrystalsPerRing, {0.0, 0.0});
  const double step = 2 * PI / h.rSectorRepeatNumber;
  const int firstOfSector0 = g.nCrystalsPerRing - g.crystalOffset;
  for (int i = 0; i < firstOfSector0; ++i)
  {
    const int k = i + g.crystalOffset;
    const double x = g.rSectorTranslationX;
    const double y = sectorY[k % perSector];
    const double angle = (k / perSector) * step;
    const double c = std::cos(angle), s = std::sin(angle);
    mXY[i] = {x * c - y * s, x * s + y * c};
  }
  for (int i = firstOfSector0; i < g.nCrystalsPerRing; ++i)
    mXY[i] = {g.rSectorTranslationX, sectorY[i - firstOfSector0]};

  const int nz = h.crystalRepeatNumbersYZ[1];
  const int mz = h.moduleRepeatNumbersYZ[1];
  const double cvz = g.crystalRepeatVectorYZ[1];
  const double mvz = g.moduleRepeatVectorYZ[1];
  const double z0 = -((mz - 1) * mvz + (nz - 1) * cvz) / 2.0;
  std::vector<double> ringZ(mz * nz);
  for (int m = 0; m < mz; ++m)
    for (int c = 0; c < nz; ++c)
      ringZ[m * nz + c] = m * mvz + c * cvz + z0;
  mZ.assign(g.nSlices, 0.0);
  for (int s = 0; s < g.nSlices; ++s)
    mZ[s] = s % 2 == 0 ? ringZ[s / 2]
                       : (ringZ[s / 2] + ringZ[(s + 1) / 2]) / 2.0;
}

void ScannerData::printContent() const
{
  echo("= Scanner header:");
  printEmptyLine();
  echo("== Crystals:");
  printVector("crystal dimensions XYZ in mm", mHeader.crystalDimsXYZ);
  printVector("crystal repeat numbers YZ", mHeader.crystalRepeatNumbersYZ);
  printVector("inter-crystal distance YZ in mm",
              mHeader.interCrystalDistanceYZ);
  printEmptyLine();
  echo("== Modules");
  printVector("module dimensions XYZ in mm", mHeader.moduleDimsXYZ);
  printVector("module repeat numbers YZ", mHeader.moduleRepeatNumbersYZ);
  printVector("inter-module distance YZ in mm", mHeader.interModuleDistanceYZ);
  printEmptyLine();
  echo("== rSectors");
  printVector("rSector dimensions XYZ in mm", mHeader.rSectorDimsXYZ);
  printValue("rSector repeat number", mHeader.rSectorRepeatNumber);
  printValue("rSector inner radius in mm", mHeader.rSectorInnerRadius);
  printEmptyLine();
  printValue("number of rings", mGeometry.nRings);
  printValue("number of crystals per ring", mGeometry.nCrystalsPerRing);
  printValue("number of slices", mGeometry.nSlices);
  printEmptyLine();
}

void ScannerData::printCrystalPositions() const
{
  for (size_t i = 0; i < mXY.size(); ++i)
    std::printf("crystal %zu: x = %.6f, y = %.6f\n", i, mXY[i].x, mXY[i].y);
  for (size_t s = 0; s < mZ.size(); ++s)
    std::printf("slice %zu: z = %.6f\n", s, mZ[s]);
}

namespace
{
// "{a,b,c}" of one field of the path elements, up to the terminator.
template<typename F>
void printPathField(const types::PathElement* path, F field)
{
  std::cout << "{";
  for (const types::PathElement* e = path; e->coord != -1; ++e)
    std::cout << (e == path ? "" : ",") << field(*e);
  std::cout << "}" << std::endl;
}
}

void ScannerData::printBinInfo(int seg, int view, int axialCoord,
                               int tangCoord, int crystal1, int slice1,
                               int crystal2, int slice2, const VolData& vol,
                               types::PathElement* path,
                               types::VoxelValue line) const
{
  printValue("seg", seg);
  printValue("view", view);
  printValue("axialCoord", axialCoord);
  printValue("tangCoord", tangCoord);
  printEmptyLine();
  printValue("crystal1", crystal1);
  printValue("slice1", slice1);
  printValue("crystal2", crystal2);
  printValue("slice2", slice2);
  printEmptyLine();
  std::cout << "X: (" << mXY[crystal1].x << "," << mXY[crystal2].x << ")"
            << std::endl;
  std::cout << "Y: (" << mXY[crystal1].y << "," << mXY[crystal2].y << ")"
            << std::endl;
  std::cout << "Z: (" << mZ[slice1] << "," << mZ[slice2] << ")" << std::endl;
  printEmptyLine();
  if (path)
  {
    const types::VoxelValue* data = vol.getDataArray();
    printPathField(path, [](const types::PathElement& e) { return e.coord; });
    printPathField(path, [](const types::PathElement& e) { return e.length; });
    std::cout << std::endl;
    printPathField(path,
                   [data](const types::PathElement& e) { return data[e.coord]; });
    std::cout << std::endl;
  }
  printValue("line", line);
  printEmptyLine();
}

void ScannerData::getCrystalCoordinates(int rSectorID, int moduleID,
                                        int crystalID, int* ring,
                                        int* crystal) const
{
  // rSectors hold moduleRepeat[0] x moduleRepeat[1] modules, each
  // crystalRepeat[0] x crystalRepeat[1] crystals (y: around the ring, z:
  // along the axis); in-ring index 0 is the middle crystal of rSector 0
  const int my = mHeader.moduleRepeatNumbersYZ[0];
  const int cy = mHeader.crystalRepeatNumbersYZ[0];
  const int cz = mHeader.crystalRepeatNumbersYZ[1];
  *ring = (moduleID / my) * cz + crystalID / cy;
  int c = (rSectorID * my + moduleID % my) * cy + crystalID % cy -
    mGeometry.crystalOffset;
  if (c >= mGeometry.nCrystalsPerRing) c -= mGeometry.nCrystalsPerRing;
  *crystal = c;
}
