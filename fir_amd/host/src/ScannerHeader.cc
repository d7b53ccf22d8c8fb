#include <ScannerHeader.h>
#include <console.h>

void ScannerHeader::setDefaults()
{
  crystalDimsXYZ.assign(3, 0.0);
  crystalRepeatNumbersYZ.assign(2, 1);
  interCrystalDistanceYZ.assign(2, 0.0);
  moduleDimsXYZ.assign(3, 0.0);
  moduleRepeatNumbersYZ.assign(2, 1);
  interModuleDistanceYZ.assign(2, 0.0);
  rSectorDimsXYZ.assign(3, 0.0);
  rSectorRepeatNumber = 0;
  rSectorInnerRadius = 0.0;
}

namespace
{
template<typename T>
bool anyNegative(const std::vector<T>& v)
{
  for (const auto& x : v)
    if (x < 0) return true;
  return false;
}

template<typename T>
bool anyNonPositive(const std::vector<T>& v)
{
  for (const auto& x : v)
    if (x <= 0) return true;
  return false;
}

// Tight size along one axis of `count` items of size `item` separated by
// `gap`.
double tight(double item, int count, double gap)
{
  return item * count + gap * (count - 1);
}

void fitOrCheck(double& value, double minimum, const char* what)
{
  if (value == 0.0)
    value = minimum;
  else if (value < minimum)
    error(what, minimum);
}
}

void ScannerHeader::check()
{
  if (crystalDimsXYZ.size() != 3)
    error("\"crystal dimensions XYZ in mm\" must contain three elements");
  if (anyNegative(crystalDimsXYZ))
    error("Elements of \"crystal dimensions XYZ in mm\" must not be "
          "negative");
  if (crystalRepeatNumbersYZ.size() != 2)
    error("\"crystal repeat numbers YZ\" must contain two elements");
  if (anyNonPositive(crystalRepeatNumbersYZ))
    error("Elements of \"crystal repeat numbers YZ\" must be > 0");
  if (interCrystalDistanceYZ.size() != 2)
    error("\"inter-crystal distance YZ in mm\" must contain two elements");
  if (anyNegative(interCrystalDistanceYZ))
    error("Elements of \"inter-crystal distance YZ in mm\" must not be "
          "negative");
  if (moduleDimsXYZ.size() != 3)
    error("\"module dimensions XYZ in mm\" must contain three elements");
  if (anyNegative(moduleDimsXYZ))
    error("Elements of \"module dimensions XYZ in mm\" must not be negative");
  if (moduleRepeatNumbersYZ.size() != 2)
    error("\"module repeat numbers YZ\" must contain two elements");
  if (anyNonPositive(moduleRepeatNumbersYZ))
    error("Elements of \"module repeat numbers YZ\" must be > 0");
  if (interModuleDistanceYZ.size() != 2)
    error("\"inter-module distance YZ in mm\" must contain two elements");
  if (anyNegative(interModuleDistanceYZ))
    error("Elements of \"inter-module distance YZ in mm\" must not be "
          "negative");
  if (rSectorDimsXYZ.size() != 3)
    error("\"rSector dimensions XYZ in mm\" must contain three elements");
  if (anyNegative(rSectorDimsXYZ))
    error("Elements of \"rSector dimensions XYZ in mm\" must not be "
          "negative");
  if (rSectorRepeatNumber <= 0)
    error("\"rSector repeat number\" must be present and > 0");
  if (rSectorInnerRadius <= 0)
    error("\"rSector inner radius in mm\" must be present and > 0");
  if (crystalRepeatNumbersYZ[0] > 1 && crystalDimsXYZ[1] == 0)
    error("Crystal repeat number in Y cannot be > 1 if crystal dimension in "
          "Y is zero");
  if (crystalRepeatNumbersYZ[1] > 1 && crystalDimsXYZ[2] == 0)
    error("Crystal repeat number in Z cannot be > 1 if crystal dimension in "
          "Z is zero");

  fitOrCheck(moduleDimsXYZ[0], crystalDimsXYZ[0],
             "Module dimension in X must be >= ");
  fitOrCheck(moduleDimsXYZ[1],
             tight(crystalDimsXYZ[1], crystalRepeatNumbersYZ[0],
                   interCrystalDistanceYZ[0]),
             "Module dimension in Y must be >= ");
  fitOrCheck(moduleDimsXYZ[2],
             tight(crystalDimsXYZ[2], crystalRepeatNumbersYZ[1],
                   interCrystalDistanceYZ[1]),
             "Module dimension in Z must be >= ");

  if (moduleRepeatNumbersYZ[0] > 1 && moduleDimsXYZ[1] == 0)
    error("Module repeat number in Y cannot be > 1 if module dimension in Y "
          "is zero");
  if (moduleRepeatNumbersYZ[1] > 1 && moduleDimsXYZ[2] == 0)
    error("Module repeat number in Z cannot be > 1 if module dimension in Z "
          "is zero");

  fitOrCheck(rSectorDimsXYZ[0], moduleDimsXYZ[0],
             "rSector dimension in X must be >= ");
  fitOrCheck(rSectorDimsXYZ[1],
             tight(moduleDimsXYZ[1], moduleRepeatNumbersYZ[0],
                   interModuleDistanceYZ[0]),
             "rSector dimension in Y must be >= ");
  fitOrCheck(rSectorDimsXYZ[2],
             tight(moduleDimsXYZ[2], moduleRepeatNumbersYZ[1],
                   interModuleDistanceYZ[1]),
             "rSector dimension in Z must be >= ");
}

void ScannerGeometry::fill(const ScannerHeader& h)
{
  crystalRepeatVectorYZ = {h.crystalDimsXYZ[1] + h.interCrystalDistanceYZ[0],
                           h.crystalDimsXYZ[2] + h.interCrystalDistanceYZ[1]};
  moduleRepeatVectorYZ = {h.moduleDimsXYZ[1] + h.interModuleDistanceYZ[0],
                          h.moduleDimsXYZ[2] + h.interModuleDistanceYZ[1]};
  rSectorTranslationX = h.rSectorInnerRadius + h.rSectorDimsXYZ[0] / 2;
  nCrystalsPerRing = h.crystalRepeatNumbersYZ[0] * h.moduleRepeatNumbersYZ[0] *
    h.rSectorRepeatNumber;
  nRings = h.crystalRepeatNumbersYZ[1] * h.moduleRepeatNumbersYZ[1];
  nCrystals = nRings * nCrystalsPerRing;
  nSlices = 2 * nRings - 1;
  crystalOffset = h.moduleRepeatNumbersYZ[0] * h.crystalRepeatNumbersYZ[0] / 2;
}
