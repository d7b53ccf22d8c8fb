// VolData.h -- multi-frame FP32 volume (x fastest) with an active frame,
// the reference's VolData interface (VolData.h:9-157). Frames are
// contiguous in one host block; element-wise arithmetic runs on the GPU.
#pragma once

#include <VolHeader.h>
#include <pinned.h>
#include <types.h>

#include <string>
#include <vector>

class VolData
{
public:
  enum class ConstructionMode
  {
    ALLOCATE,             // allocate without initialising
    INITIALIZE,           // set every voxel to initValue
    READ_DATA,            // read the data file (copy: copy the values)
    READ_DATA_IF_PROVIDED // read if a data file is named, else initialise
  };

  VolData();
  VolData(const std::string& inputVolFile,
          ConstructionMode mode = ConstructionMode::READ_DATA,
          types::VoxelValue initValue = 0.0);
  VolData(const VolData& vol,
          ConstructionMode mode = ConstructionMode::READ_DATA,
          types::VoxelValue initValue = 0.0);
  explicit VolData(const VolHeader& header);
  VolData(VolData&& other) noexcept;
  ~VolData() = default;

  VolData& operator=(const VolData&) = delete;
  VolData& operator=(VolData&& other) noexcept;

  void read(const std::string& headerFileName,
            ConstructionMode mode = ConstructionMode::READ_DATA,
            types::VoxelValue initValue = 0.0);
  void copy(const VolData& vol,
            ConstructionMode mode = ConstructionMode::READ_DATA,
            types::VoxelValue initValue = 0.0);
  void write(const std::string& outputVolFile) const;

  void allocateAsMultiVol(const VolData& templateVol, int nFrames);
  void allocateSingleFrameFromMultiVol(VolData& templateVol);

  void printContent() const;
  void checkNFrames(int nFrames) const;

  void assign(const VolData& vol);
  void assignFrame(const VolData& vol, int frame);
  void setAllVoxels(types::VoxelValue value = 0.0);
  void setAllVoxelsAllFrames(types::VoxelValue value = 0.0);

  // a = (a > 1e-5 && b > 1e-5) ? a op b : 0 on the active frames (GPU)
  VolData& operator*=(const VolData& inputVol);
  VolData& operator/=(const VolData& inputVol);

  // Line integral of the active frame along a path (elements up to the
  // terminator coord == -1), accumulated in path order into a float as
  // float(line + length * value): VolData::computeLineIntegral
  // (VolData.cc:357-371). Host-side, for single paths; the projectors run
  // on the GPU.
  types::VoxelValue computeLineIntegral(
    types::PathElement* pathElementsArray) const;
  // Adds length * line to every voxel of the path in the active frame, each
  // update atomic (safe from concurrent threads):
  // VolData::projectLineIntegral (VolData.cc:373-385).
  void projectLineIntegral(types::PathElement* pathElementsArray,
                           types::VoxelValue line);

  void setActiveFrame(int frame) const;
  int getActiveFrame() const { return mActiveFrame; }

  const VolHeader& getHeader() const { return mHeader; }
  int getNFrames() const { return mHeader.nFrames; }
  const types::VolExtent& getVolExtent() const { return mGeometry.volExtent; }
  int getNVoxelsPerFrame() const { return mGeometry.nVoxelsPerFrame; }
  bool isAllocated() const { return mDataArray != nullptr; }
  types::VoxelValue* getDataArray() const { return mDataArray; }
  // Start of frame 0 (all frames, contiguous).
  types::VoxelValue* getAllFrames() const
  {
    return mFrames.empty() ? nullptr : mFrames[0];
  }

  void setVoxel(int i, int j, int k, types::VoxelValue value);
  types::VoxelValue getVoxel(int i, int j, int k) const;

private:
  void allocate();
  void deallocate();
  void rebindFrames();

  VolHeader mHeader{};
  VolGeometry mGeometry{};
  mutable int mActiveFrame = 0;
  mutable types::VoxelValue* mDataArray = nullptr;
  std::vector<types::VoxelValue*> mFrames;
  // page-locked when a GPU is present (direct DMA to and from the device)
  std::vector<types::VoxelValue, pinned::Allocator<types::VoxelValue>>
    mStorage;
};
