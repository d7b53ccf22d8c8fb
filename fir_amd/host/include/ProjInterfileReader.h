// ProjInterfileReader.h -- .hs/.s sinogram reader and writer (reference
// ProjInterfileReader.h:8-35). Data: raw native FP32, segments concatenated
// from -segOffset to +segOffset, each [view][axial][tang].
#pragma once

#include <ProjHeader.h>
#include <streamio.h>
#include <types.h>

#include <string>

class ProjInterfileReader
{
public:
  explicit ProjInterfileReader(const std::string& headerFileName);

  ProjHeader getHeader() const { return mHeader; }
  ProjGeometry getGeometry() const { return mGeometry; }

  // Reads nBins floats into the segment arrays, several chunks at a time.
  // When the segment arrays are one contiguous block in file order (as
  // ProjData keeps them), `landed` runs after each chunk of it has been
  // read (offsets in bytes from the first bin), so that device copies of
  // the chunks overlap the rest of the read.
  void readData(types::BinValue** dataArray,
                const streamio::ChunkHook& landed = {});

  // Writes the header and the segment arrays; when they are one contiguous
  // block, `ready` runs before each chunk of it is written, in order.
  static void writeProjInterfile(const std::string& outputProjFile,
                                 const ProjHeader& header,
                                 types::BinValue** dataArray,
                                 const streamio::ChunkHook& ready = {});

private:
  std::string mDataFileName;
  ProjHeader mHeader;
  ProjGeometry mGeometry;
};
