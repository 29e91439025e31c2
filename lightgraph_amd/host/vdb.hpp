// Reader of OpenVDB FloatGrids (the .vdb files the reference loads with
// openvdb::io::File, main.cc:486-502) into a dense array for DenseBrickGrid.
//
// Written from OpenVDB's published file layout (file format versions 222 and
// later, OpenVDB 3 onwards): archive header, grid descriptors with stream
// offsets, per-grid compression flags, metadata, transform, the
// Tree_float_5_4_3 topology (root tiles and children, 32^3 and 16^3 internal
// nodes with child and value masks and mask-compressed tile values, 8^3 leaf
// value masks) and the leaf buffers (mask-compressed values, zlib or blosc, or
// half floats).  No OpenVDB code is used; parity with OpenVDB itself is
// unpinned (the reference ships no .vdb file to check against).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace lightgraph {

struct VdbFloatGrid {
    std::string name;
    double voxel_size = 1.0;         // uniform voxel size (world units)
    double translation[3] = {0, 0, 0};  // world position of index (0, 0, 0)
    float background = 0.0f;
    bool empty = true;               // no active voxel
    int32_t active_min[3] = {0, 0, 0}, active_max[3] = {-1, -1, -1};  // inclusive, index space
    // dense values (z-y-x order) over index box [dense_min, dense_min + dims)
    int32_t dense_min[3] = {0, 0, 0}, dims[3] = {0, 0, 0};
    std::vector<float> dense;
    uint64_t leaves = 0, tiles = 0;  // nodes read
};

// Names of the grids in a file (suffixes of duplicate names stripped).
std::vector<std::string> vdbGridNames(const std::string &path);

// Reads FloatGrid `name`.  The dense box is the active voxel bounding box
// (OpenVDB's evalActiveVoxelBoundingBox), or [box_min, box_min + box_dims) when
// given (the temperature grid is read over the density's box).  Throws
// std::runtime_error on a missing grid, another value type, an unsupported map
// or a malformed stream.
VdbFloatGrid readVdbFloatGrid(const std::string &path, const std::string &name, const int32_t *box_min = nullptr,
                              const int32_t *box_dims = nullptr);

// The grids the renderer reads from a file (main.cc:488-502): "density" (required)
// and, when present, "temperature" on the density's voxels.  A temperature grid
// with the density's transform is read over the density's box as stored; one
// with another transform is trilinearly resampled (OpenVDB's BoxSampler) at the
// density's voxel centres, background outside its stored values.
struct VdbVolume {
    VdbFloatGrid density;
    bool has_temperature = false;
    bool temperature_resampled = false;
    std::vector<float> temperature;  // density.dims, z-y-x
};
VdbVolume readVdbVolume(const std::string &path);

}  // namespace lightgraph
