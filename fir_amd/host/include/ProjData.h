// ProjData.h -- sinogram container, the reference's ProjData interface
// (ProjData.h:23-170). Storage is one contiguous FP32 block in file order
// (segments -S..S, each [view][axial][tang]) with per-segment pointers into
// it, so getDataArray()[seg + segOffset][bin] works as in the reference and
// one copy moves the whole sinogram to the GPU.
#pragma once

#include <ProjHeader.h>
#include <pinned.h>
#include <types.h>

#include <string>
#include <tuple>
#include <vector>

class ProjData
{
public:
  enum class ConstructionMode
  {
    ALLOCATE,
    INITIALIZE,
    READ_DATA
  };

  ProjData();
  ProjData(const std::string& inputProjFile,
           ConstructionMode mode = ConstructionMode::READ_DATA,
           types::BinValue initValue = 0.0);
  ProjData(const ProjData& proj,
           ConstructionMode mode = ConstructionMode::READ_DATA,
           types::BinValue initValue = 0.0);
  ProjData(ProjData&& other) noexcept;
  ProjData& operator=(ProjData&& other) noexcept;
  ProjData& operator=(const ProjData&) = delete;
  ~ProjData() = default;

  void read(const std::string& headerFileName,
            ConstructionMode mode = ConstructionMode::READ_DATA,
            types::BinValue initValue = 0.0);
  void copy(const ProjData& proj,
            ConstructionMode mode = ConstructionMode::READ_DATA,
            types::BinValue initValue = 0.0);
  void write(const std::string& outputProjFile) const;

  void printContent() const;
  void checkNSubsets(int nSubsets) const;
  void setAllBins(types::BinValue value = 0.0);

  // {crystalAxialCoord1, crystalAxialCoord2}: slice indices of the crystals
  // of (seg, axialCoord) (reference ProjData.cc:198-299).
  std::tuple<int, int> getCrystalAxialCoord(int seg, int axialCoord) const;
  // {crystalAngCoord1, crystalAngCoord2} of (view, tangCoord)
  // (reference ProjData.cc:301-340).
  std::tuple<int, int> getCrystalAngCoord(int view, int tangCoord) const;

  // Sinogram coordinates of the LOR between crystal (crystalAxialCoord1 =
  // ring, crystalAngCoord1) and crystal (ring2, crystal2); false if the
  // sinogram does not hold it (reference ProjData.cc:342-479).
  bool getBinCoordinates(int crystalAxialCoord1, int crystalAngCoord1,
                         int crystalAxialCoord2, int crystalAngCoord2,
                         int* seg, int* view, int* axialCoord,
                         int* tangCoord) const;

  // Span > 1: weight every bin by 1 / (ring pairs fused into its
  // (segment, axial) position) where that is above 1 (reference
  // ProjData.cc:541-615), on the GPU.
  void rebinWeight();

  // Clamped product and ACF exponential, run on the GPU.
  ProjData& operator*=(const ProjData& inputProj);
  void exponential();

  const ProjHeader& getHeader() const { return mHeader; }
  const ProjGeometry& getGeometry() const { return mGeometry; }
  types::BinValue** getDataArray() const
  {
    return mSegments.empty() ? nullptr
                             : const_cast<types::BinValue**>(mSegments.data());
  }
  // All bins, contiguous, in file order.
  types::BinValue* getContiguousData() const
  {
    return mStorage.empty() ? nullptr
                            : const_cast<types::BinValue*>(mStorage.data());
  }

  void setBin(int seg, int view, int axialCoord, int tangCoord,
              types::BinValue value);
  types::BinValue getBin(int seg, int view, int axialCoord,
                         int tangCoord) const;
  void incrementBin(int seg, int view, int axialCoord, int tangCoord);
  void weightBin(int seg, int view, int axialCoord, int tangCoord,
                 types::BinValue weight);

private:
  long long index(int seg, int view, int axialCoord, int tangCoord) const;
  void allocate(bool initialize, types::BinValue initValue);

  ProjHeader mHeader{};
  ProjGeometry mGeometry{};
  // page-locked when a GPU is present (direct DMA to and from the device)
  std::vector<types::BinValue, pinned::Allocator<types::BinValue>> mStorage;
  std::vector<types::BinValue*> mSegments;
};
