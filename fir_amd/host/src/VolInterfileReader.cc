#include <KeyParser.h>
#include <VolInterfileReader.h>
#include <console.h>
#include <tools.h>
#include <writeKeys.h>

#include <cstdint>
#include <filesystem>
#include <fstream>

namespace
{
const char* const kShortFloat = "SHORT FLOAT";
const char* const kLongFloat = "LONG FLOAT";
const char* const kFloat = "FLOAT";
const char* const kUnsigned = "UNSIGNED INTEGER";
const char* const kSigned = "SIGNED INTEGER";
const char* const kLittle = "LITTLEENDIAN";
const char* const kBig = "BIGENDIAN";

bool needSwap(DataByteOrderEnum order)
{
  return systemIsLittleEndian() ? order != DataByteOrderEnum::LITTLEENDIAN
                                : order != DataByteOrderEnum::BIGENDIAN;
}

template<typename T>
void readFrames(const VolInterfileReaderParams& p, const VolHeader& h,
                const VolGeometry& g, std::vector<types::VoxelValue*>& frames)
{
  std::ifstream in(p.dataFileName, std::ios::binary);
  if (!in.is_open()) error("Couldn't open file ", p.dataFileName);
  in.seekg(p.dataOffset, std::ios::beg);
  const bool swap = needSwap(p.dataByteOrder);
  std::vector<T> buf(g.nVoxelsPerFrame);
  for (int f = 0; f < h.nFrames; ++f)
  {
    in.read(reinterpret_cast<char*>(buf.data()),
            static_cast<std::streamsize>(buf.size() * sizeof(T)));
    if (!in)
      error("The number of pixels that were read from the data file (",
            in.gcount(), ") is inferior to the number expected from the ",
            "header file (", g.nVoxelsTotal, ")");
    for (int i = 0; i < g.nVoxelsPerFrame; ++i)
      frames[f][i] = static_cast<types::VoxelValue>(
        swap ? swapEndianness(buf[i]) : buf[i]);
  }
}
}

void VolInterfileReaderParams::setDefaults()
{
  dataFileName.clear();
  dataOffset = 0;
  dataTypeAsString.clear();
  dataType = DataTypeEnum::NONE;
  bytesPerPixel = 0;
  dataByteOrderAsString.clear();
  dataByteOrder = DataByteOrderEnum::NONE;
}

void VolInterfileReaderParams::check()
{
  if (dataOffset < 0) error("Data offset must not be negative");
  dataTypeAsString = strToUpper(dataTypeAsString);
  dataByteOrderAsString = strToUpper(dataByteOrderAsString);

  int fallback = 0;
  if (dataTypeAsString == kShortFloat)
  {
    dataType = DataTypeEnum::FLOAT;
    if (bytesPerPixel != 4) fallback = 4;
  }
  else if (dataTypeAsString == kLongFloat)
  {
    dataType = DataTypeEnum::FLOAT;
    if (bytesPerPixel != 8) fallback = 8;
  }
  else if (dataTypeAsString == kFloat)
  {
    dataType = DataTypeEnum::FLOAT;
    if (bytesPerPixel != 4 && bytesPerPixel != 8) fallback = 4;
  }
  else if (dataTypeAsString == kUnsigned || dataTypeAsString == kSigned)
  {
    dataType = dataTypeAsString == kUnsigned ? DataTypeEnum::UNSIGNED_INTEGER
                                             : DataTypeEnum::SIGNED_INTEGER;
    if (bytesPerPixel != 1 && bytesPerPixel != 2 && bytesPerPixel != 4 &&
        bytesPerPixel != 8)
      fallback = 4;
  }
  else
    error("Unrecognized data type name ", dataTypeAsString);

  if (fallback)
  {
    bytesPerPixel = fallback;
    warning("Invalid value for bytes per pixel. Using default for datatype ",
            dataTypeAsString, ", which is ", bytesPerPixel, ".");
  }

  if (dataByteOrderAsString == kBig)
    dataByteOrder = DataByteOrderEnum::BIGENDIAN;
  else if (dataByteOrderAsString == kLittle)
    dataByteOrder = DataByteOrderEnum::LITTLEENDIAN;
  else if (dataByteOrderAsString.empty())
  {
    const bool little = systemIsLittleEndian();
    dataByteOrderAsString = little ? kLittle : kBig;
    dataByteOrder =
      little ? DataByteOrderEnum::LITTLEENDIAN : DataByteOrderEnum::BIGENDIAN;
    warning("Invalid value for imagedata byte order. Using default for "
            "current machine (",
            kLittle, ").");
  }
  else
    error("Unrecognized byte order", dataByteOrderAsString);
}

VolInterfileReader::VolInterfileReader(const std::string& headerFileName)
{
  mParams.setDefaults();
  mHeader.setDefaults();

  KeyParser kp;
  kp.addStartKey("!INTERFILE");
  kp.addKey("name of data file", &mParams.dataFileName);
  kp.addKey("data offset in bytes", &mParams.dataOffset);
  kp.addKey("number format", &mParams.dataTypeAsString);
  kp.addKey("number of bytes per pixel", &mParams.bytesPerPixel);
  kp.addKey("imagedata byte order", &mParams.dataByteOrderAsString);
  kp.addKey("matrix size [1]", &mHeader.volSize.nPixelsX);
  kp.addKey("matrix size [2]", &mHeader.volSize.nPixelsY);
  kp.addKey("number of slices", &mHeader.volSize.nSlices);
  kp.addKey("scaling factor (mm/pixel) [1]", &mHeader.voxelExtent.pixelWidth);
  kp.addKey("scaling factor (mm/pixel) [2]", &mHeader.voxelExtent.pixelHeight);
  kp.addKey("slice thickness (pixels)", &mHeader.voxelExtent.sliceThickness);
  kp.addKey("first pixel offset (mm) [1]", &mHeader.volOffset.x);
  kp.addKey("first pixel offset (mm) [2]", &mHeader.volOffset.y);
  kp.addKey("first pixel offset (mm) [3]", &mHeader.volOffset.z);
  kp.addKey("number of time frames", &mHeader.nFrames);
  kp.addStopKey("!END OF INTERFILE");
  kp.parse(headerFileName);

  addPath(headerFileName, mParams.dataFileName);
  mParams.check();
  mHeader.check();
  mGeometry.fill(mHeader);
}

void VolInterfileReader::readData(std::vector<types::VoxelValue*>& frames)
{
  const int b = mParams.bytesPerPixel;
  switch (mParams.dataType)
  {
  case DataTypeEnum::UNSIGNED_INTEGER:
    if (b == 1) readFrames<std::uint8_t>(mParams, mHeader, mGeometry, frames);
    if (b == 2) readFrames<std::uint16_t>(mParams, mHeader, mGeometry, frames);
    if (b == 4) readFrames<std::uint32_t>(mParams, mHeader, mGeometry, frames);
    if (b == 8) readFrames<std::uint64_t>(mParams, mHeader, mGeometry, frames);
    break;
  case DataTypeEnum::SIGNED_INTEGER:
    if (b == 1) readFrames<std::int8_t>(mParams, mHeader, mGeometry, frames);
    if (b == 2) readFrames<std::int16_t>(mParams, mHeader, mGeometry, frames);
    if (b == 4) readFrames<std::int32_t>(mParams, mHeader, mGeometry, frames);
    if (b == 8) readFrames<std::int64_t>(mParams, mHeader, mGeometry, frames);
    break;
  case DataTypeEnum::FLOAT:
    if (b == 4) readFrames<float>(mParams, mHeader, mGeometry, frames);
    if (b == 8) readFrames<double>(mParams, mHeader, mGeometry, frames);
    break;
  default:
    error("Invalid value for data type enum");
  }
}

void VolInterfileReader::writeVolInterfile(
  const std::string& outputVolFile,
  const VolHeader& h,
  const std::vector<types::VoxelValue*>& frames)
{
  std::filesystem::path headerPath(outputVolFile);
  headerPath.replace_extension(".h33");
  std::filesystem::path dataPath(outputVolFile);
  dataPath.replace_extension(".i33");

  VolInterfileReaderParams p;
  p.setDefaults();
  p.dataFileName = dataPath.filename().string();
  p.dataTypeAsString = kFloat;
  p.bytesPerPixel = 4;
  p.dataByteOrderAsString = kLittle;
  p.check();

  {
    std::ofstream os(headerPath);
    if (!os.is_open()) error("Couldn't create file ", headerPath.string());
    writeKey(os, "!INTERFILE");
    os << '\n';
    writeKey(os, "name of data file", p.dataFileName);
    writeKey(os, "data offset in bytes", p.dataOffset);
    writeKey(os, "number format", p.dataTypeAsString);
    writeKey(os, "number of bytes per pixel", p.bytesPerPixel);
    writeKey(os, "imagedata byte order", p.dataByteOrderAsString);
    os << '\n';
    writeKey(os, "number of dimensions", 3);
    os << '\n';
    writeKey(os, "matrix size [1]", h.volSize.nPixelsX);
    writeKey(os, "matrix size [2]", h.volSize.nPixelsY);
    writeKey(os, "matrix size [3]", h.volSize.nSlices);
    writeKey(os, "number of slices", h.volSize.nSlices);
    os << '\n';
    writeKey(os, "scaling factor (mm/pixel) [1]", h.voxelExtent.pixelWidth);
    writeKey(os, "scaling factor (mm/pixel) [2]", h.voxelExtent.pixelHeight);
    writeKey(os, "scaling factor (mm/pixel) [3]", h.voxelExtent.sliceThickness);
    writeKey(os, "slice thickness (pixels)", h.voxelExtent.sliceThickness);
    os << '\n';
    writeKey(os, "first pixel offset (mm) [1]", h.volOffset.x);
    writeKey(os, "first pixel offset (mm) [2]", h.volOffset.y);
    writeKey(os, "first pixel offset (mm) [3]", h.volOffset.z);
    os << '\n';
    writeKey(os, "number of time frames", h.nFrames);
    os << '\n';
    writeKey(os, "!END OF INTERFILE");
  }

  std::ofstream data(dataPath, std::ios::binary);
  if (!data.is_open()) error("Couldn't create file ", dataPath.string());
  const int n = h.volSize.nPixelsX * h.volSize.nPixelsY * h.volSize.nSlices;
  const bool swap = needSwap(p.dataByteOrder);
  std::vector<float> buf(n);
  for (int f = 0; f < h.nFrames; ++f)
  {
    for (int i = 0; i < n; ++i)
      buf[i] = swap ? swapEndianness(frames[f][i]) : frames[f][i];
    data.write(reinterpret_cast<const char*>(buf.data()),
               static_cast<std::streamsize>(n * sizeof(float)));
  }
}
