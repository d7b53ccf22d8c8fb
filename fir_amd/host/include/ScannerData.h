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

private:
  void computeTables();

  ScannerHeader mHeader{};
  ScannerGeometry mGeometry{};
  std::vector<types::SpatialCoords2D> mXY;
  std::vector<types::SpatialCoord> mZ;
};
