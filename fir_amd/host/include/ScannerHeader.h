// ScannerHeader.h -- cylindrical scanner description (reference
// ScannerHeader.h:7-68): crystals grouped in modules grouped in rSectors,
// rSectors repeated around the axis.
#pragma once

#include <types.h>

#include <vector>

struct ScannerHeader
{
  std::vector<types::SpatialCoord> crystalDimsXYZ;
  std::vector<int> crystalRepeatNumbersYZ;
  std::vector<types::SpatialCoord> interCrystalDistanceYZ;
  std::vector<types::SpatialCoord> moduleDimsXYZ;
  std::vector<int> moduleRepeatNumbersYZ;
  std::vector<types::SpatialCoord> interModuleDistanceYZ;
  std::vector<types::SpatialCoord> rSectorDimsXYZ;
  int rSectorRepeatNumber;
  types::SpatialCoord rSectorInnerRadius;

  void setDefaults();
  // Validates and replaces zero module/rSector dimensions by the tightest
  // enclosing size.
  void check();
};

struct ScannerGeometry
{
  std::vector<types::SpatialCoord> crystalRepeatVectorYZ;
  std::vector<types::SpatialCoord> moduleRepeatVectorYZ;
  types::SpatialCoord rSectorTranslationX;
  int nRings, nCrystalsPerRing, nCrystals;
  int nSlices;
  int crystalOffset;

  void fill(const ScannerHeader& header);
};
