// ProjInterfileReader.h -- .hs/.s sinogram reader and writer (reference
// ProjInterfileReader.h:8-35). Data: raw native FP32, segments concatenated
// from -segOffset to +segOffset, each [view][axial][tang].
#pragma once

#include <ProjHeader.h>
#include <types.h>

#include <string>

class ProjInterfileReader
{
public:
  explicit ProjInterfileReader(const std::string& headerFileName);

  ProjHeader getHeader() const { return mHeader; }
  ProjGeometry getGeometry() const { return mGeometry; }

  // Reads nBins floats into the segment arrays.
  void readData(types::BinValue** dataArray);

  static void writeProjInterfile(const std::string& outputProjFile,
                                 const ProjHeader& header,
                                 types::BinValue** dataArray);

private:
  std::string mDataFileName;
  ProjHeader mHeader;
  ProjGeometry mGeometry;
};
