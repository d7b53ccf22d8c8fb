// VolInterfileReader.h -- .h33/.i33 volume reader and writer (reference
// VolInterfileReader.h:7-86). Reads unsigned/signed integers of 1-8 bytes
// and floats of 4/8 bytes in either byte order; writes FLOAT, 4 bytes,
// LITTLEENDIAN.
#pragma once

#include <VolHeader.h>

#include <string>
#include <vector>

enum class DataTypeEnum
{
  NONE,
  UNSIGNED_INTEGER,
  SIGNED_INTEGER,
  FLOAT
};

enum class DataByteOrderEnum
{
  NONE,
  LITTLEENDIAN,
  BIGENDIAN
};

struct VolInterfileReaderParams
{
  std::string dataFileName;
  int dataOffset;
  std::string dataTypeAsString;
  DataTypeEnum dataType;
  int bytesPerPixel;
  std::string dataByteOrderAsString;
  DataByteOrderEnum dataByteOrder;

  void setDefaults();
  void check();
};

class VolInterfileReader
{
public:
  explicit VolInterfileReader(const std::string& headerFileName);

  bool hasDataFile() const { return !mParams.dataFileName.empty(); }
  VolHeader getHeader() const { return mHeader; }
  VolGeometry getGeometry() const { return mGeometry; }

  void readData(std::vector<types::VoxelValue*>& frameVector);

  static void writeVolInterfile(
    const std::string& outputVolFile,
    const VolHeader& header,
    const std::vector<types::VoxelValue*>& frameVector);

private:
  VolInterfileReaderParams mParams;
  VolHeader mHeader;
  VolGeometry mGeometry;
};
