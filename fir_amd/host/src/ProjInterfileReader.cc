#include <KeyParser.h>
#include <ProjInterfileReader.h>
#include <console.h>
#include <tools.h>
#include <writeKeys.h>

#include <filesystem>
#include <fstream>
#include <vector>

ProjInterfileReader::ProjInterfileReader(const std::string& headerFileName)
{
  mHeader.setDefaults();
  KeyParser kp;
  kp.addStartKey("!PROJECTION DATA PARAMETERS");
  kp.addKey("name of data file", &mDataFileName);
  kp.addKey("number of rings", &mHeader.nRings);
  kp.addKey("number of crystals per ring", &mHeader.nCrystalsPerRing);
  kp.addKey("segment span", &mHeader.segmentSpan);
  kp.addKey("number of segments", &mHeader.nSegments);
  kp.addKey("number of tangential coordinates", &mHeader.nTangCoords);
  kp.addStopKey("!END OF PROJECTION DATA PARAMETERS");
  kp.parse(headerFileName);
  addPath(headerFileName, mDataFileName);
  mHeader.check();
  mGeometry.fill(mHeader);
}

void ProjInterfileReader::readData(types::BinValue** data)
{
  std::ifstream in(mDataFileName, std::ios::binary);
  if (!in.is_open()) error("Couldn't open file ", mDataFileName);
  for (int seg = -mGeometry.segOffset; seg <= mGeometry.segOffset; ++seg)
  {
    const long long n = static_cast<long long>(mGeometry.nViews) *
      mGeometry.getNAxialCoords(seg) * mHeader.nTangCoords;
    in.read(reinterpret_cast<char*>(data[seg + mGeometry.segOffset]),
            static_cast<std::streamsize>(n * sizeof(types::BinValue)));
    if (!in)
      error("The number of bins that were read from the data file (",
            in.gcount(), ") is inferior to that expected from the header "
            "file (",
            mGeometry.nBins, ")");
  }
}

void ProjInterfileReader::writeProjInterfile(const std::string& outputProjFile,
                                             const ProjHeader& h,
                                             types::BinValue** data)
{
  ProjGeometry g;
  g.fill(h);
  std::filesystem::path headerPath(outputProjFile);
  headerPath.replace_extension(".hs");
  std::filesystem::path dataPath(outputProjFile);
  dataPath.replace_extension(".s");
  {
    std::ofstream os(headerPath);
    if (!os.is_open()) error("Couldn't create file ", headerPath.string());
    writeKey(os, "!PROJECTION DATA PARAMETERS");
    writeKey(os, "name of data file", dataPath.filename().string());
    writeKey(os, "number of rings", h.nRings);
    writeKey(os, "number of crystals per ring", h.nCrystalsPerRing);
    writeKey(os, "segment span", h.segmentSpan);
    writeKey(os, "number of segments", h.nSegments);
    writeKey(os, "number of tangential coordinates", h.nTangCoords);
    writeKey(os, "!END OF PROJECTION DATA PARAMETERS");
  }
  std::ofstream os(dataPath, std::ios::binary);
  if (!os.is_open()) error("Couldn't create file ", dataPath.string());
  for (int seg = -g.segOffset; seg <= g.segOffset; ++seg)
  {
    const long long n =
      static_cast<long long>(g.nViews) * g.getNAxialCoords(seg) * h.nTangCoords;
    os.write(reinterpret_cast<const char*>(data[seg + g.segOffset]),
             static_cast<std::streamsize>(n * sizeof(types::BinValue)));
  }
}
