// VolHeader.h -- volume header and derived geometry (reference
// VolHeader.h:6-52): voxel counts, voxel extents, offsets, frame count.
#pragma once

#include <types.h>

struct VolHeader
{
  types::VolSize volSize;
  types::VoxelExtent voxelExtent;
  // x, y: centre of the first voxel; z: offset of the volume centre
  types::SpatialCoords3D volOffset;
  int nFrames;

  // Frame counts are not compared.
  bool operator==(const VolHeader& o) const
  {
    return volSize == o.volSize && voxelExtent == o.voxelExtent &&
      volOffset == o.volOffset;
  }
  bool operator!=(const VolHeader& o) const { return !(*this == o); }

  void setDefaults();
  void check() const;
};

struct VolGeometry
{
  types::VolExtent volExtent;
  int nVoxelsPerFrame;
  int nVoxelsTotal;

  void fill(const VolHeader& header);
};
