// ScannerData.h -- scanner geometry and crystal position tables, the
// reference's ScannerData interface (ScannerData.h:11-53). Crystal (x, y)
// is tabulated per angular index and z per slice (2*nRings - 1 slices; odd
// slices lie midway between rings).
#pragma once

#include <ScannerHeader.h>
#include <types.h>

#include <string>
#include <vector>

class ProjData;
class VolData;

class ScannerData
{
public:
  ScannerData() = default;
  explicit ScannerData(const std::string& scannerFile);
  explicit ScannerData(const ScannerHeader& header);

  // Throws unless the sinogram has this scanner's ring and crystal counts.
  void checkProjData(const ProjData& proj) const;

  const types::SpatialCoords2D* getCrystalXYPositionVector() const
  {
    return mXY.data();
  }
  const types::SpatialCoord* getSliceZPositionVector() const
  {
    return mZ.data();
  }

  const ScannerHeader& getHeader() const { return mHeader; }
  const ScannerGeometry& getGeometry() const { return mGeometry; }

  void printContent() const;
  void printCrystalPositions() const;
  // Bin, crystal pair, crystal positions and, if given, the path (voxel
  // indices, lengths, the values of vol along it) and its line integral:
  // ScannerData::printBinInfo (ScannerData.cc:113-200), same text.
  void printBinInfo(int seg, int view, int axialCoord, int tangCoord,
                    int crystal1, int slice1, int crystal2, int slice2,
                    const VolData& vol,
                    types::PathElement* pathElementsArray = 0,
                    types::VoxelValue line = 0) const;

  // Ring and in-ring crystal coordinates of crystal crystalID of module
  // moduleID of rSector rSectorID: ScannerData::getCrystalCoordinates
  // (ScannerData.cc:202-231).
  void getCrystalCoordinates(int rSectorID, int moduleID, int crystalID,
                             int* ring, int* crystal) const;

private:
  void computeTables();

  ScannerHeader mHeader{};
  ScannerGeometry mGeometry{};
  std::vector<types::SpatialCoords2D> mXY;
  std::vector<types::SpatialCoord> mZ;
};
