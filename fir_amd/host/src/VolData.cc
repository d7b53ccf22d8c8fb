#include <VolData.h>
#include <VolInterfileReader.h>
#include <console.h>
#include <device.h>

#include <algorithm>
#include <cstring>

VolData::VolData() = default;

VolData::VolData(const std::string& file, ConstructionMode mode,
                 types::VoxelValue init)
{
  read(file, mode, init);
}

VolData::VolData(const VolData& vol, ConstructionMode mode,
                 types::VoxelValue init)
{
  copy(vol, mode, init);
}

VolData::VolData(const VolHeader& header)
{
  header.check();
  mHeader = header;
  mGeometry.fill(mHeader);
  allocate();
  setAllVoxelsAllFrames(0.0f);
}

VolData::VolData(VolData&& o) noexcept { *this = std::move(o); }

VolData& VolData::operator=(VolData&& o) noexcept
{
  if (this == &o) return *this;
  mHeader = o.mHeader;
  mGeometry = o.mGeometry;
  mStorage = std::move(o.mStorage);
  const int active = o.mActiveFrame;
  o.deallocate();
  rebindFrames();
  if (!mFrames.empty())
  {
    mActiveFrame = active;
    mDataArray = mFrames[active];
  }
  return *this;
}

void VolData::rebindFrames()
{
  mFrames.clear();
  mActiveFrame = 0;
  mDataArray = nullptr;
  if (mStorage.empty()) return;
  for (int f = 0; f < mHeader.nFrames; ++f)
    mFrames.push_back(mStorage.data() +
                      static_cast<size_t>(f) * mGeometry.nVoxelsPerFrame);
  mDataArray = mFrames[0];
}

void VolData::allocate()
{
  mStorage.assign(static_cast<size_t>(mGeometry.nVoxelsTotal), 0.0f);
  rebindFrames();
}

void VolData::deallocate()
{
  mStorage.clear();
  mStorage.shrink_to_fit();
  mFrames.clear();
  mActiveFrame = 0;
  mDataArray = nullptr;
}

void VolData::read(const std::string& file, ConstructionMode mode,
                   types::VoxelValue init)
{
  deallocate();
  VolInterfileReader reader(file);
  mHeader = reader.getHeader();
  mGeometry = reader.getGeometry();
  allocate();
  switch (mode)
  {
  case ConstructionMode::ALLOCATE:
    break;
  case ConstructionMode::INITIALIZE:
    setAllVoxelsAllFrames(init);
    break;
  case ConstructionMode::READ_DATA:
    reader.readData(mFrames);
    break;
  case ConstructionMode::READ_DATA_IF_PROVIDED:
    if (reader.hasDataFile())
      reader.readData(mFrames);
    else
      setAllVoxelsAllFrames(init);
    break;
  }
}

void VolData::copy(const VolData& vol, ConstructionMode mode,
                   types::VoxelValue init)
{
  deallocate();
  mHeader = vol.mHeader;
  mGeometry = vol.mGeometry;
  allocate();
  switch (mode)
  {
  case ConstructionMode::ALLOCATE:
    break;
  case ConstructionMode::INITIALIZE:
    setAllVoxelsAllFrames(init);
    break;
  case ConstructionMode::READ_DATA:
    if (!vol.isAllocated())
      error("Mode is READ_DATA and volume to copy isn't allocated");
    assign(vol);
    break;
  case ConstructionMode::READ_DATA_IF_PROVIDED:
    if (vol.isAllocated())
      assign(vol);
    else
      setAllVoxelsAllFrames(init);
    break;
  }
}

void VolData::write(const std::string& file) const
{
  if (!isAllocated()) error("Volume data is not allocated");
  VolInterfileReader::writeVolInterfile(file, mHeader, mFrames);
}

void VolData::allocateAsMultiVol(const VolData& tmpl, int nFrames)
{
  if (nFrames <= 0) error("Number of frames must be greater than zero");
  deallocate();
  mHeader = tmpl.mHeader;
  mGeometry = tmpl.mGeometry;
  mHeader.nFrames = nFrames;
  mGeometry.nVoxelsTotal = mGeometry.nVoxelsPerFrame * nFrames;
  allocate();
}

void VolData::allocateSingleFrameFromMultiVol(VolData& tmpl)
{
  deallocate();
  mHeader = tmpl.mHeader;
  mGeometry = tmpl.mGeometry;
  mHeader.nFrames = 1;
  mGeometry.nVoxelsTotal = mGeometry.nVoxelsPerFrame;
  allocate();
}

void VolData::printContent() const
{
  echo("= Volume header:");
  printEmptyLine();
  echo("== Volume dimensions (width in x, height in y, depth in z):");
  printVector<types::Size>("volume size in voxels",
                           {mHeader.volSize.nPixelsX, mHeader.volSize.nPixelsY,
                            mHeader.volSize.nSlices});
  printVector<types::SpatialExtent>(
    "voxel extent in mm",
    {mHeader.voxelExtent.pixelWidth, mHeader.voxelExtent.pixelHeight,
     mHeader.voxelExtent.sliceThickness});
  printVector<types::SpatialExtent>(
    "volume extent in mm",
    {mGeometry.volExtent.sliceWidth, mGeometry.volExtent.sliceHeight,
     mGeometry.volExtent.volDepth});
  printEmptyLine();
  echo("== Volume position:");
  printVector<types::SpatialCoord>("x,y coordinates of first voxel in mm",
                                   {mHeader.volOffset.x, mHeader.volOffset.y});
  printValue("volume offset in z in mm", mHeader.volOffset.z);
  printEmptyLine();
  echo("== Number of frames and voxels:");
  printValue("number of time frames", mHeader.nFrames);
  printValue("number of voxels per frame", mGeometry.nVoxelsPerFrame);
  printValue("total number of voxels", mGeometry.nVoxelsTotal);
  printEmptyLine();
}

void VolData::checkNFrames(int nFrames) const
{
  if (mHeader.nFrames != nFrames)
    error("Wrong number of frames (expected ", nFrames, ")");
}

void VolData::assign(const VolData& vol)
{
  if (!isAllocated()) error("Volume not allocated");
  if (mHeader != vol.mHeader || mHeader.nFrames != vol.mHeader.nFrames)
    error("Volume assigned is not the right size");
  std::memcpy(mFrames[0], vol.mFrames[0],
              sizeof(float) * static_cast<size_t>(mGeometry.nVoxelsTotal));
}

void VolData::assignFrame(const VolData& vol, int frame)
{
  if (!isAllocated()) error("Volume not allocated");
  if (mHeader != vol.mHeader) error("Volume assigned is not the right size");
  std::memcpy(mDataArray, vol.mFrames[frame],
              sizeof(float) * static_cast<size_t>(mGeometry.nVoxelsPerFrame));
}

void VolData::setAllVoxels(types::VoxelValue value)
{
  if (!isAllocated()) error("Volume not allocated");
  std::fill(mDataArray, mDataArray + mGeometry.nVoxelsPerFrame, value);
}

void VolData::setAllVoxelsAllFrames(types::VoxelValue value)
{
  if (!isAllocated()) error("Volume not allocated");
  std::fill(mStorage.begin(), mStorage.end(), value);
}

VolData& VolData::operator*=(const VolData& in)
{
  device::clampedOp(device::ClampedOp::Mul, mDataArray, in.mDataArray,
                    mGeometry.nVoxelsPerFrame);
  return *this;
}

VolData& VolData::operator/=(const VolData& in)
{
  device::clampedOp(device::ClampedOp::Div, mDataArray, in.mDataArray,
                    mGeometry.nVoxelsPerFrame);
  return *this;
}

types::VoxelValue VolData::computeLineIntegral(
  types::PathElement* path) const
{
  types::VoxelValue line = 0.0f;
  for (const types::PathElement* e = path; e->coord != -1; ++e)
    line = static_cast<types::VoxelValue>(
      line + e->length * mDataArray[e->coord]);
  return line;
}

void VolData::projectLineIntegral(types::PathElement* path,
                                  types::VoxelValue line)
{
  for (const types::PathElement* e = path; e->coord != -1; ++e)
  {
    types::VoxelValue* v = mDataArray + e->coord;
    types::VoxelValue seen;
    __atomic_load(v, &seen, __ATOMIC_RELAXED);
    types::VoxelValue next;
    do
      next = static_cast<types::VoxelValue>(seen + e->length * line);
    while (!__atomic_compare_exchange(v, &seen, &next, false,
                                      __ATOMIC_RELAXED, __ATOMIC_RELAXED));
  }
}

void VolData::setActiveFrame(int frame) const
{
  if (!(frame >= 0 && frame < mHeader.nFrames)) error("Invalid frame");
  mActiveFrame = frame;
  mDataArray = mFrames[frame];
}

void VolData::setVoxel(int i, int j, int k, types::VoxelValue value)
{
  mDataArray[i + j * mHeader.volSize.nPixelsX +
             k * mHeader.volSize.nPixelsX * mHeader.volSize.nPixelsY] = value;
}

types::VoxelValue VolData::getVoxel(int i, int j, int k) const
{
  return mDataArray[i + j * mHeader.volSize.nPixelsX +
                    k * mHeader.volSize.nPixelsX * mHeader.volSize.nPixelsY];
}
