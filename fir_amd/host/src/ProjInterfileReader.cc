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

namespace
{
// Bins of each segment, and whether the segment arrays are one contiguous
// block in file order.
bool contiguous(const ProjGeometry& g, int nTang, types::BinValue** data,
                std::vector<long long>* counts)
{
  counts->clear();
  bool one = true;
  for (int seg = -g.segOffset; seg <= g.segOffset; ++seg)
  {
    const long long n =
      static_cast<long long>(g.nViews) * g.getNAxialCoords(seg) * nTang;
    const int k = seg + g.segOffset;
    if (k > 0 && data[k] != data[k - 1] + counts->back()) one = false;
    counts->push_back(n);
  }
  return one;
}

void shortRead(long long got, long long expected)
{
  error("The number of bins that were read from the data file (", got,
        ") is inferior to that expected from the header file (", expected,
        ")");
}
}

void ProjInterfileReader::readData(types::BinValue** data,
                                   const streamio::ChunkHook& landed)
{
  std::vector<long long> counts;
  const long long nBins = mGeometry.nBins;
  const size_t fb = sizeof(types::BinValue);
  if (contiguous(mGeometry, mHeader.nTangCoords, data, &counts))
  {
    const size_t got =
      streamio::readFile(mDataFileName, 0, data[0], nBins * fb, landed);
    if (got < nBins * fb) shortRead(static_cast<long long>(got / fb), nBins);
    return;
  }
  size_t at = 0;
  for (size_t k = 0; k < counts.size(); ++k)
  {
    const size_t bytes = counts[k] * fb;
    const size_t got = streamio::readFile(mDataFileName, at, data[k], bytes);
    if (got < bytes) shortRead(static_cast<long long>((at + got) / fb), nBins);
    at += bytes;
  }
}

void ProjInterfileReader::writeProjInterfile(const std::string& outputProjFile,
                                             const ProjHeader& h,
                                             types::BinValue** data,
                                             const streamio::ChunkHook& ready)
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
  std::vector<long long> counts;
  const size_t fb = sizeof(types::BinValue);
  if (contiguous(g, h.nTangCoords, data, &counts))
  {
    streamio::writeFile(dataPath.string(), data[0], g.nBins * fb, ready);
    return;
  }
  std::ofstream os(dataPath, std::ios::binary);
  if (!os.is_open()) error("Couldn't create file ", dataPath.string());
  for (size_t k = 0; k < counts.size(); ++k)
    os.write(reinterpret_cast<const char*>(data[k]),
             static_cast<std::streamsize>(counts[k] * fb));
}
