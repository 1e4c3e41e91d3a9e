// Tile files: a portable on-disk form of an octree of 8^3 tiles (levels,
// coordinates, tile types, cell types and any of the tile channels), and a
// reader for the reference's raw HAHostTileHolder dump
// (src/PoissonIOFunc.cu:38-124), which writes the tile structs as they lie in
// memory and is therefore tied to its compiler's struct layout.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace cg {

// In-memory content of a tile file: tiles in file order.
struct TileSet {
  double h0 = 0.0;
  std::vector<int32_t> level;      // [n]
  std::vector<int32_t> coord;      // [3n]
  std::vector<uint8_t> type;       // [n]
  std::vector<uint8_t> ctype;      // [512 n]
  std::vector<int32_t> channel_id; // [m]
  std::vector<float> channel;      // [m][512 n], i-major cells
  int64_t size() const { return int64_t(level.size()); }
};

// Portable format, version 1 (all fields little-endian):
//   0   char[8]  magic "CIRTILES"
//   8   uint32   version (1)
//   12  uint32   number of channels m
//   16  int64    number of tiles n
//   24  float64  h0 (level-0 cell size)
//   32  uint64   FNV-1a 64 checksum of every byte from offset 64 to the end
//   40  byte[24] zero
//   64  int32 level[n], int32 coord[3n], uint8 type[n], uint8 cell_type[512 n],
//       then per channel: int32 channel id (0..10, the reference's channel
//       numbers), int32 element type (1 = float32), float32 data[512 n]
// Tile types are LEAF 1, NONLEAF 2, GHOST 4; cell types DIRICHLET 1,
// INTERIOR 2, NEUMANN 4; cells of a tile are i-major (i*64 + j*8 + k).
void write_tile_file(const std::string& path, const TileSet& s);
TileSet read_tile_file(const std::string& path);

// The reference's HAHostTileHolder<PoissonTile<float>> dump as written by
// WriteHAHostTileHolderToFile on x86-64: the tiles that the level lists point
// to, with their cell types and the requested channels.
TileSet read_holder_dump(const std::string& path, const std::vector<int32_t>& channels);

}  // namespace cg
