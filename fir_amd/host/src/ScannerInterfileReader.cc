#include <KeyParser.h>
#include <ScannerInterfileReader.h>

void readScannerInterfileHeader(const std::string& file, ScannerHeader& h)
{
  h.setDefaults();
  KeyParser kp;
  kp.addStartKey("!SCANNER PARAMETERS");
  kp.addKey("crystal dimensions XYZ in mm", &h.crystalDimsXYZ);
  kp.addKey("crystal repeat numbers YZ", &h.crystalRepeatNumbersYZ);
  kp.addKey("inter-crystal distance YZ in mm", &h.interCrystalDistanceYZ);
  kp.addKey("module dimensions XYZ in mm", &h.moduleDimsXYZ);
  kp.addKey("module repeat numbers YZ", &h.moduleRepeatNumbersYZ);
  kp.addKey("inter-module distance YZ in mm", &h.interModuleDistanceYZ);
  kp.addKey("rSector dimensions XYZ in mm", &h.rSectorDimsXYZ);
  kp.addKey("rSector repeat number", &h.rSectorRepeatNumber);
  kp.addKey("rSector inner radius in mm", &h.rSectorInnerRadius);
  kp.addStopKey("!END OF SCANNER PARAMETERS");
  kp.parse(file);
  h.check();
}
