// Tile files (portable format) and the reader of the reference's raw
// HAHostTileHolder dump.
#include "tilefile.hpp"

#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>

namespace cg {

namespace {

constexpr char kMagic[8] = {'C', 'I', 'R', 'T', 'I', 'L', 'E', 'S'};
constexpr uint32_t kVersion = 1;
constexpr int64_t kHeader = 64;
constexpr int kCells = 512;
constexpr int kRefChannels = 11;  // PoissonTile::num_channels (src/PoissonTile.h:23)

struct FileCloser {
  void operator()(FILE* f) const { if (f) std::fclose(f); }
};
using File = std::unique_ptr<FILE, FileCloser>;

File open_file(const std::string& path, const char* mode) {
  File f(std::fopen(path.c_str(), mode));
  if (!f) throw std::runtime_error("cannot open " + path);
  return f;
}

int64_t file_size(FILE* f) {
  if (std::fseek(f, 0, SEEK_END) != 0) throw std::runtime_error("seek failed");
  const long n = std::ftell(f);
  if (n < 0) throw std::runtime_error("tell failed");
  std::rewind(f);
  return int64_t(n);
}

// FNV-1a, 64 bit
struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void* p, size_t n) {
    const uint8_t* b = static_cast<const uint8_t*>(p);
    for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 1099511628211ull; }
  }
};

bool valid_tile_type(uint8_t t) { return t == 1 || t == 2 || t == 4; }
bool valid_cell_type(uint8_t t) { return t == 1 || t == 2 || t == 4; }

void check_set(const TileSet& s) {
  const size_t n = s.level.size();
  if (s.coord.size() != 3 * n || s.type.size() != n || s.ctype.size() != size_t(kCells) * n ||
      s.channel.size() != s.channel_id.size() * size_t(kCells) * n)
    throw std::runtime_error("tile set arrays of inconsistent sizes");
  for (size_t t = 0; t < n; t++) {
    if (s.level[t] < 0 || s.level[t] > 30) throw std::runtime_error("tile level out of range");
    if (!valid_tile_type(s.type[t])) throw std::runtime_error("invalid tile type");
  }
  for (uint8_t c : s.ctype) if (!valid_cell_type(c)) throw std::runtime_error("invalid cell type");
  for (int32_t id : s.channel_id)
    if (id < 0 || id >= kRefChannels) throw std::runtime_error("channel id out of range");
}

template <class T>
void put(FILE* f, Fnv* h, const T* p, size_t n) {
  if (n && std::fwrite(p, sizeof(T), n, f) != n) throw std::runtime_error("write failed");
  if (h) h->add(p, sizeof(T) * n);
}
template <class T>
void get(FILE* f, T* p, size_t n) {
  if (n && std::fread(p, sizeof(T), n, f) != n) throw std::runtime_error("file truncated");
}

}  // namespace

void write_tile_file(const std::string& path, const TileSet& s) {
  check_set(s);
  const size_t n = s.level.size(), m = s.channel_id.size();
  File f = open_file(path, "wb");
  uint8_t head[kHeader] = {};
  put(f.get(), nullptr, head, sizeof(head));  // rewritten below with the checksum
  Fnv h;
  put(f.get(), &h, s.level.data(), n);
  put(f.get(), &h, s.coord.data(), 3 * n);
  put(f.get(), &h, s.type.data(), n);
  put(f.get(), &h, s.ctype.data(), kCells * n);
  for (size_t c = 0; c < m; c++) {
    const int32_t tag[2] = {s.channel_id[c], 1};
    put(f.get(), &h, tag, 2);
    put(f.get(), &h, s.channel.data() + c * kCells * n, kCells * n);
  }
  std::memcpy(head, kMagic, 8);
  const uint32_t ver = kVersion, nch = uint32_t(m);
  const int64_t nt = int64_t(n);
  std::memcpy(head + 8, &ver, 4);
  std::memcpy(head + 12, &nch, 4);
  std::memcpy(head + 16, &nt, 8);
  std::memcpy(head + 24, &s.h0, 8);
  std::memcpy(head + 32, &h.h, 8);
  std::rewind(f.get());
  put(f.get(), nullptr, head, sizeof(head));
  if (std::fflush(f.get()) != 0) throw std::runtime_error("write failed");
}

TileSet read_tile_file(const std::string& path) {
  File f = open_file(path, "rb");
  const int64_t bytes = file_size(f.get());
  if (bytes < kHeader) throw std::runtime_error(path + ": not a tile file (too short)");
  uint8_t head[kHeader];
  get(f.get(), head, sizeof(head));
  if (std::memcmp(head, kMagic, 8) != 0) throw std::runtime_error(path + ": not a tile file (bad magic)");
  uint32_t ver, nch;
  int64_t n;
  uint64_t sum;
  TileSet s;
  std::memcpy(&ver, head + 8, 4);
  std::memcpy(&nch, head + 12, 4);
  std::memcpy(&n, head + 16, 8);
  std::memcpy(&s.h0, head + 24, 8);
  std::memcpy(&sum, head + 32, 8);
  if (ver != kVersion) throw std::runtime_error(path + ": unsupported tile file version");
  if (n < 0 || n > (int64_t(1) << 31) || nch > uint32_t(kRefChannels))
    throw std::runtime_error(path + ": bad tile or channel count");
  const int64_t expect = kHeader + n * (4 + 12 + 1 + kCells) + int64_t(nch) * (8 + n * kCells * 4);
  if (bytes != expect) throw std::runtime_error(path + ": size does not match the header (truncated or corrupt)");
  s.level.resize(n);
  s.coord.resize(3 * n);
  s.type.resize(n);
  s.ctype.resize(size_t(kCells) * n);
  s.channel_id.resize(nch);
  s.channel.resize(size_t(nch) * kCells * n);
  get(f.get(), s.level.data(), n);
  get(f.get(), s.coord.data(), 3 * n);
  get(f.get(), s.type.data(), n);
  get(f.get(), s.ctype.data(), kCells * n);
  Fnv h;
  h.add(s.level.data(), 4 * n);
  h.add(s.coord.data(), 12 * n);
  h.add(s.type.data(), n);
  h.add(s.ctype.data(), kCells * n);
  for (uint32_t c = 0; c < nch; c++) {
    int32_t tag[2];
    get(f.get(), tag, 2);
    if (tag[1] != 1) throw std::runtime_error(path + ": unsupported channel element type");
    s.channel_id[c] = tag[0];
    float* dst = s.channel.data() + size_t(c) * kCells * n;
    get(f.get(), dst, size_t(kCells) * n);
    h.add(tag, 8);
    h.add(dst, size_t(kCells) * n * 4);
  }
  if (h.h != sum) throw std::runtime_error(path + ": checksum mismatch");
  check_set(s);
  return s;
}

// The reference's dump (WriteHAHostTileHolderToFile, src/PoissonIOFunc.cu:38-75):
//   float mH0, int mNumLayers, int mMaxLevel (src/HAGrid.h:80-82);
//   size_t number of tiles, then each PoissonTile<float> as raw bytes;
//   size_t number of levels, per level a size_t count and per entry the
//   HATileInfo's Coord (3 x int32), int mLevel, uint8 mType, then an intptr_t
//   index into the tile array (-1: none), each written unpadded.
// PoissonTile<float> (src/PoissonTile.h:50-56) on x86-64: float mData[11][736]
// at 0, HATileInfo mNeighbors[6] (32 bytes each: Coord, pointer, int, uint8,
// src/HATileInfo.h:69-72) at 32384, uint8 mCellType[512] at 32576, two bools,
// int mStatus, int mSerialIdx; 33104 bytes with the 8-byte alignment the
// pointers impose.
TileSet read_holder_dump(const std::string& path, const std::vector<int32_t>& channels) {
  constexpr int64_t kTileBytes = 33104, kChnl = 736, kCtypeOff = 32576;
  constexpr int64_t kEntry = 12 + 4 + 1 + 8;
  for (int32_t id : channels)
    if (id < 0 || id >= kRefChannels) throw std::runtime_error("channel id out of range");
  File f = open_file(path, "rb");
  const int64_t bytes = file_size(f.get());
  float h0f;
  int32_t layers, max_level;
  uint64_t ntiles;
  if (bytes < 20) throw std::runtime_error(path + ": too short for a tile holder dump");
  get(f.get(), &h0f, 1);
  get(f.get(), &layers, 1);
  get(f.get(), &max_level, 1);
  get(f.get(), &ntiles, 1);
  if (!(h0f > 0) || layers <= 0 || layers > 31 || ntiles > (uint64_t(1) << 31) ||
      int64_t(20 + ntiles * kTileBytes) > bytes)
    throw std::runtime_error(path + ": not a tile holder dump of PoissonTile<float> (header)");
  const int64_t tiles_at = 20;
  if (std::fseek(f.get(), long(tiles_at + int64_t(ntiles) * kTileBytes), SEEK_SET) != 0)
    throw std::runtime_error("seek failed");
  uint64_t nlev;
  get(f.get(), &nlev, 1);
  if (nlev > 31) throw std::runtime_error(path + ": bad level count");
  struct Entry { int32_t c[3]; int32_t level; uint8_t type; int64_t index; };
  std::vector<Entry> entries;
  int64_t pos = tiles_at + int64_t(ntiles) * kTileBytes + 8;
  for (uint64_t l = 0; l < nlev; l++) {
    uint64_t cnt;
    get(f.get(), &cnt, 1);
    pos += 8;
    if (pos + int64_t(cnt) * kEntry > bytes) throw std::runtime_error(path + ": level list truncated");
    for (uint64_t e = 0; e < cnt; e++) {
      uint8_t raw[kEntry];
      get(f.get(), raw, kEntry);
      Entry E;
      std::memcpy(E.c, raw, 12);
      std::memcpy(&E.level, raw + 12, 4);
      E.type = raw[16];
      std::memcpy(&E.index, raw + 17, 8);
      if (E.index >= 0) {
        if (uint64_t(E.index) >= ntiles) throw std::runtime_error(path + ": tile index out of range");
        if (E.level != int32_t(l)) throw std::runtime_error(path + ": entry level differs from its list");
        if (!valid_tile_type(E.type)) throw std::runtime_error(path + ": invalid tile type");
        entries.push_back(E);
      }
    }
    pos += int64_t(cnt) * kEntry;
  }
  if (pos != bytes) throw std::runtime_error(path + ": trailing or missing bytes after the level lists");
  TileSet s;
  s.h0 = double(h0f);
  const size_t n = entries.size();
  s.level.resize(n);
  s.coord.resize(3 * n);
  s.type.resize(n);
  s.ctype.resize(size_t(kCells) * n);
  s.channel_id = channels;
  s.channel.resize(channels.size() * size_t(kCells) * n);
  std::vector<uint8_t> tile(kTileBytes);
  for (size_t t = 0; t < n; t++) {
    const Entry& E = entries[t];
    s.level[t] = E.level;
    for (int d = 0; d < 3; d++) s.coord[3 * t + d] = E.c[d];
    s.type[t] = E.type;
    if (std::fseek(f.get(), long(tiles_at + E.index * kTileBytes), SEEK_SET) != 0) throw std::runtime_error("seek failed");
    get(f.get(), tile.data(), size_t(kTileBytes));
    std::memcpy(&s.ctype[size_t(kCells) * t], tile.data() + kCtypeOff, kCells);
    for (size_t c = 0; c < channels.size(); c++)
      std::memcpy(&s.channel[(c * n + t) * kCells], tile.data() + size_t(channels[c]) * kChnl * 4, kCells * 4);
  }
  check_set(s);
  return s;
}

}  // namespace cg
