// types.h -- scalar and small aggregate types of the FIR host API.
// Same names and meanings as the reference (types.h:1-79): int sizes and
// indices, double spatial coordinates, float voxel and bin values.
#pragma once

namespace types
{
using Size = int;
using Index = Size;

using SpatialExtent = double;
using SpatialCoord = SpatialExtent;

using VoxelValue = float;
using BinValue = float;

struct VolSize
{
  Size nPixelsX;
  Size nPixelsY;
  Size nSlices;

  Size nVoxelsTotal() const { return nPixelsX * nPixelsY * nSlices; }
  bool operator==(const VolSize& o) const
  {
    return nPixelsX == o.nPixelsX && nPixelsY == o.nPixelsY &&
      nSlices == o.nSlices;
  }
};

struct VoxelExtent
{
  SpatialExtent pixelWidth;
  SpatialExtent pixelHeight;
  SpatialExtent sliceThickness;

  bool operator==(const VoxelExtent& o) const
  {
    return pixelWidth == o.pixelWidth && pixelHeight == o.pixelHeight &&
      sliceThickness == o.sliceThickness;
  }
};

struct VolExtent
{
  SpatialExtent sliceWidth;
  SpatialExtent sliceHeight;
  SpatialExtent volDepth;
};

inline VolExtent operator*(const VolSize& s, const VoxelExtent& e)
{
  return {s.nPixelsX * e.pixelWidth, s.nPixelsY * e.pixelHeight,
          s.nSlices * e.sliceThickness};
}

struct SpatialCoords3D
{
  SpatialCoord x;
  SpatialCoord y;
  SpatialCoord z;

  bool operator==(const SpatialCoords3D& o) const
  {
    return x == o.x && y == o.y && z == o.z;
  }
};

struct SpatialCoords2D
{
  SpatialCoord x;
  SpatialCoord y;
};

// One element of a Siddon path: linear voxel index and length in mm.
struct PathElement
{
  Index coord;
  SpatialExtent length;
};
}
