#include <VolHeader.h>
#include <console.h>

void VolHeader::setDefaults()
{
  volSize = {0, 0, 0};
  voxelExtent = {0.0, 0.0, 0.0};
  volOffset = {0.0, 0.0, 0.0};
  nFrames = 1;
}

void VolHeader::check() const
{
  if (volSize.nPixelsX <= 0 || volSize.nPixelsY <= 0 || volSize.nSlices <= 0)
    error("Number of voxels must be greater than zero in each dimension");
  if (voxelExtent.pixelWidth <= 0 || voxelExtent.pixelHeight <= 0 ||
      voxelExtent.sliceThickness <= 0)
    error("Voxel extent must be greater than zero in each dimension");
  if (nFrames <= 0) error("Number of frames must be greater than zero");
}

void VolGeometry::fill(const VolHeader& h)
{
  volExtent = h.volSize * h.voxelExtent;
  nVoxelsPerFrame = h.volSize.nPixelsX * h.volSize.nPixelsY * h.volSize.nSlices;
  nVoxelsTotal = nVoxelsPerFrame * h.nFrames;
}
