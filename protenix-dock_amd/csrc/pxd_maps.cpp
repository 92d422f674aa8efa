// Map directories: save and restore of the grid maps (receptor_cache::save /
// restore, engine/src/core/grid.cpp:305-431) in this library's own format,
// which pxdock.h specifies byte for byte.  Everything here is host code; the
// device is touched only through pxd_grid_download / pxd_grid_upload.
#include <errno.h>
#include <sys/stat.h>
#include <sys/types.h>
#include <unistd.h>

#include <algorithm>
#include <cinttypes>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "pxd_internal.hpp"
#include "pxdock.h"
#include "pxdock_vdw_table.h"

extern "C" pxd_status pxd_set_error_(pxd_status st, const char* msg);

namespace {

#define TAG_ROW(i, tag, s, e) tag,
const char* const kTypeTag[PXD_NUM_VDW_TYPES] = {PXD_VDW_TABLE(TAG_ROW)};
#undef TAG_ROW

const char kMagic[8] = {'P', 'X', 'D', 'M', 'A', 'P', 0, 1};
const char* const kMetaName = "meta.pxd";
const char* const kMetaTmp = "meta.pxd.tmp";
constexpr size_t kHeaderBytes = 64;
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull;
constexpr uint64_t kFnvPrime = 0x100000001b3ull;

pxd_status err(pxd_status st, const char* fmt, ...) {
  char buf[768];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return pxd_set_error_(st, buf);
}

struct Fnv {
  uint64_t h = kFnvBasis;
  void bytes(const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
      h ^= c[i];
      h *= kFnvPrime;
    }
  }
  template <class T>
  void value(T v) {
    bytes(&v, sizeof(v));  // x86-64 and gfx950 hosts are little-endian
  }
};

// FNV-1a 64 over little-endian 64-bit words (n_bytes is a multiple of 8)
uint64_t payload_checksum(const void* p, size_t n_bytes, uint64_t h = kFnvBasis) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i + 8 <= n_bytes; i += 8) {
    uint64_t w;
    memcpy(&w, c + i, 8);
    h ^= w;
    h *= kFnvPrime;
  }
  return h;
}

std::string join(const std::string& dir, const std::string& name) {
  if (dir.empty() || dir.back() == '/') return dir + name;
  return dir + "/" + name;
}

std::string map_file_name(int type) {
  return std::string(type < 0 ? "coul" : kTypeTag[type]) + ".pxm";
}

bool is_dir(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

pxd_status make_dirs(const std::string& dir) {
  if (dir.empty()) return err(PXD_ERR_ARG, "maps: empty directory name");
  std::string cur;
  size_t pos = 0;
  while (pos <= dir.size()) {
    size_t next = dir.find('/', pos);
    if (next == std::string::npos) next = dir.size();
    cur = dir.substr(0, next);
    if (!cur.empty() && !is_dir(cur) && mkdir(cur.c_str(), 0777) != 0 && errno != EEXIST)
      return err(PXD_ERR_IO, "maps: cannot create directory %s: %s", cur.c_str(), strerror(errno));
    pos = next + 1;
  }
  if (!is_dir(dir)) return err(PXD_ERR_IO, "maps: %s is not a directory", dir.c_str());
  return PXD_OK;
}

long long n_points(const pxd_box& b) {
  return (long long)b.npoints[0] * b.npoints[1] * b.npoints[2];
}

pxd_status check_meta(const pxd_maps_meta& m) {
  for (int d = 0; d < 3; ++d)
    if (m.box.npoints[d] < 2) return err(PXD_ERR_CONF, "maps: npoints[%d] = %d < 2", d, m.box.npoints[d]);
  if (!(m.box.spacing > 0)) return err(PXD_ERR_CONF, "maps: spacing must be > 0");
  if (m.n_types < 0 || m.n_types > PXD_NUM_VDW_TYPES)
    return err(PXD_ERR_CONF, "maps: %d types", m.n_types);
  for (int t = 0; t < m.n_types; ++t) {
    if (m.types[t] < 0 || m.types[t] >= PXD_NUM_VDW_TYPES)
      return err(PXD_ERR_CONF, "maps: type %d out of range", m.types[t]);
    if (t > 0 && m.types[t] <= m.types[t - 1])
      return err(PXD_ERR_CONF, "maps: types must be ascending and distinct");
  }
  if (m.n_atoms < 0 || m.n_excluded < 0 || m.n_excluded > m.n_atoms)
    return err(PXD_ERR_CONF, "maps: %d excluded of %d receptor atoms", m.n_excluded, m.n_atoms);
  return PXD_OK;
}

// derived box fields as decide_grid_dims / fill_derived_fields (grid.cpp:57-86)
void fill_derived(pxd_box* b) {
  b->spacing_inv = 1.0 / b->spacing;
  for (int d = 0; d < 3; ++d) {
    b->npoints_minus_one[d] = b->npoints[d] - 1;
    b->oppo_xyz[d] = b->init_xyz[d] + b->npoints_minus_one[d] * b->spacing;
  }
}

struct MapHeader {
  char magic[8];
  uint32_t version;
  int32_t type;
  int32_t npoints[3];
  int32_t channels;
  uint64_t payload_bytes;
  uint64_t checksum;
  char pad[16];
};
static_assert(sizeof(MapHeader) == kHeaderBytes, "map header is 64 bytes");

pxd_status write_map(const std::string& path, int type, const pxd_box& box, const float* data,
                     uint64_t* checksum) {
  const long long P = n_points(box);
  MapHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, kMagic, 8);
  h.version = PXD_MAPS_VERSION;
  h.type = type;
  for (int d = 0; d < 3; ++d) h.npoints[d] = box.npoints[d];
  h.channels = 4;
  h.payload_bytes = (uint64_t)P * 16;
  h.checksum = payload_checksum(data, h.payload_bytes);
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return err(PXD_ERR_IO, "maps: cannot write %s: %s", path.c_str(), strerror(errno));
  bool ok = fwrite(&h, sizeof(h), 1, f) == 1 &&
            fwrite(data, 1, h.payload_bytes, f) == h.payload_bytes;
  ok = (fclose(f) == 0) && ok;
  if (!ok) return err(PXD_ERR_IO, "maps: short write to %s", path.c_str());
  *checksum = h.checksum;
  return PXD_OK;
}

// Reads one map file into dst (P * 4 floats), checking the header against the
// metadata and the payload against both checksums.
pxd_status read_map(const std::string& path, int type, const pxd_box& box, uint64_t meta_sum,
                    float* dst) {
  const char* tag = type < 0 ? "coul" : kTypeTag[type];
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return err(PXD_ERR_IO, "Map file for atom type [%s] is missing! (%s)", tag, path.c_str());
  MapHeader h;
  const uint64_t want = (uint64_t)n_points(box) * 16;
  pxd_status st = PXD_OK;
  if (fread(&h, sizeof(h), 1, f) != 1) {
    st = err(PXD_ERR_IO, "maps: %s is truncated (no header)", path.c_str());
  } else if (memcmp(h.magic, kMagic, 8) != 0) {
    st = err(PXD_ERR_CONF, "maps: %s is not a map file", path.c_str());
  } else if (h.version != PXD_MAPS_VERSION) {
    st = err(PXD_ERR_CONF, "Format version of map file %s is not compatible! Expected [%d] vs. Got [%u]",
             path.c_str(), PXD_MAPS_VERSION, h.version);
  } else if (h.type != type || h.channels != 4 || h.npoints[0] != box.npoints[0] ||
             h.npoints[1] != box.npoints[1] || h.npoints[2] != box.npoints[2] ||
             h.payload_bytes != want) {
    st = err(PXD_ERR_CONF, "maps: header of %s disagrees with the metadata", path.c_str());
  } else if (h.checksum != meta_sum) {
    st = err(PXD_ERR_CONF, "maps: checksum of %s disagrees with the metadata", path.c_str());
  } else if (fread(dst, 1, want, f) != want) {
    st = err(PXD_ERR_IO, "maps: %s is truncated", path.c_str());
  } else if (fgetc(f) != EOF) {
    st = err(PXD_ERR_IO, "maps: %s has trailing bytes", path.c_str());
  } else if (payload_checksum(dst, want) != h.checksum) {
    st = err(PXD_ERR_IO, "maps: %s fails its checksum", path.c_str());
  }
  fclose(f);
  return st;
}

struct Meta {
  pxd_maps_meta m;
  std::vector<int32_t> excluded;
  uint64_t sum[PXD_NUM_VDW_TYPES + 1];  // [0] Coulomb, [1 + t] type slot t
};

pxd_status parse_meta(const std::string& dir, Meta* out) {
  const std::string path = join(dir, kMetaName);
  std::ifstream in(path);
  if (!in) return err(PXD_ERR_IO, "The metadata file of maps is missing! (%s)", path.c_str());
  Meta& M = *out;
  memset(&M.m, 0, sizeof(M.m));
  M.excluded.clear();
  std::string line, key;
  int lineno = 0;
  auto bad = [&](const char* what) {
    return err(PXD_ERR_CONF, "maps: %s line %d: %s", path.c_str(), lineno, what);
  };
  // records in the order pxdock.h specifies
  const char* order[] = {"pxdock-maps", "init_xyz", "npoints", "spacing", "slope",
                         "receptor",    "excluded", "types"};
  std::vector<bool> have_map;
  for (int rec = 0; rec < 8; ++rec) {
    if (!std::getline(in, line)) return bad("file ends early");
    ++lineno;
    std::istringstream ls(line);
    if (!(ls >> key) || key != order[rec]) return bad("unexpected record");
    bool ok = true;
    switch (rec) {
      case 0: {
        long long v;
        ok = bool(ls >> v);
        if (ok && v != PXD_MAPS_VERSION)
          return err(PXD_ERR_CONF,
                     "Format version of map files are not compatible! Expected [%d] vs. Got [%lld]",
                     PXD_MAPS_VERSION, v);
        break;
      }
      case 1:
        ok = bool(ls >> M.m.box.init_xyz[0] >> M.m.box.init_xyz[1] >> M.m.box.init_xyz[2]);
        break;
      case 2:
        ok = bool(ls >> M.m.box.npoints[0] >> M.m.box.npoints[1] >> M.m.box.npoints[2]);
        break;
      case 3: ok = bool(ls >> M.m.box.spacing); break;
      case 4: ok = bool(ls >> M.m.slope); break;
      case 5: {
        std::string hex;
        ok = bool(ls >> M.m.n_atoms >> hex) && hex.size() == 16;
        if (ok) {
          char* end = nullptr;
          M.m.receptor_key = strtoull(hex.c_str(), &end, 16);
          ok = end && *end == 0;
        }
        break;
      }
      case 6: {
        ok = bool(ls >> M.m.n_excluded) && M.m.n_excluded >= 0 && M.m.n_excluded <= M.m.n_atoms;
        for (int i = 0; ok && i < M.m.n_excluded; ++i) {
          int32_t a;
          ok = bool(ls >> a) && a >= 0 && a < M.m.n_atoms &&
               (M.excluded.empty() || a > M.excluded.back());
          if (ok) M.excluded.push_back(a);
        }
        break;
      }
      case 7: {
        ok = bool(ls >> M.m.n_types) && M.m.n_types >= 0 && M.m.n_types <= PXD_NUM_VDW_TYPES;
        for (int i = 0; ok && i < M.m.n_types; ++i) ok = bool(ls >> M.m.types[i]);
        break;
      }
    }
    std::string extra;
    if (!ok) return bad("malformed values");
    if (ls >> extra) return bad("trailing values");
  }
  pxd_status st = check_meta(M.m);
  if (st != PXD_OK) return st;
  fill_derived(&M.m.box);
  have_map.assign(M.m.n_types + 1, false);
  bool ended = false;
  while (std::getline(in, line)) {
    ++lineno;
    std::istringstream ls(line);
    if (!(ls >> key)) continue;
    if (key == "end") {
      ended = true;
      break;
    }
    std::string which, file, hex, extra;
    if (key != "map" || !(ls >> which >> file >> hex) || hex.size() != 16 || (ls >> extra))
      return bad("malformed map record");
    int slot = -1;
    if (which == "coul") {
      slot = 0;
    } else {
      char* end = nullptr;
      long t = strtol(which.c_str(), &end, 10);
      if (!end || *end) return bad("malformed map type");
      for (int i = 0; i < M.m.n_types; ++i)
        if (M.m.types[i] == t) slot = i + 1;
    }
    if (slot < 0) return bad("map of a type not listed");
    if (have_map[slot]) return bad("map listed twice");
    if (file != map_file_name(slot == 0 ? -1 : M.m.types[slot - 1])) return bad("unexpected map file name");
    char* end = nullptr;
    M.sum[slot] = strtoull(hex.c_str(), &end, 16);
    if (!end || *end) return bad("malformed checksum");
    have_map[slot] = true;
  }
  if (!ended) return bad("no end record (incomplete metadata)");
  for (int s = 0; s <= M.m.n_types; ++s)
    if (!have_map[s]) return bad("a listed map has no map record");
  return PXD_OK;
}

}  // namespace

extern "C" {

const char* pxd_vdw_type_tag(int32_t t) {
  if (t >= 0 && t < PXD_NUM_VDW_TYPES) return kTypeTag[t];
  return t == PXD_NUM_VDW_TYPES ? "coul" : nullptr;
}

pxd_status pxd_receptor_key(const pxd_mol* mol, const double* xyz, const pxd_leaves* lv,
                            uint64_t* key) {
  if (!mol || !xyz || !key || !mol->charge || !mol->sigma || !mol->epsilon)
    return err(PXD_ERR_ARG, "receptor_key: null argument");
  const int n = mol->n_atoms;
  std::vector<char> movable(n > 0 ? n : 0, 0);
  if (lv)
    for (int k = 0; k < lv->n_leaves; ++k)
      for (int a = lv->atom_off[k]; a < lv->atom_off[k + 1]; ++a)
        if (lv->atoms[a] >= 0 && lv->atoms[a] < n) movable[lv->atoms[a]] = 1;
  Fnv f;
  f.value<int32_t>(n);
  for (int i = 0; i < n; ++i) {
    if (movable[i]) continue;
    f.value<int32_t>(i);
    f.value<double>(xyz[3 * i]);
    f.value<double>(xyz[3 * i + 1]);
    f.value<double>(xyz[3 * i + 2]);
    f.value<double>(mol->charge[i]);
    f.value<double>(mol->sigma[i]);
    f.value<double>(mol->epsilon[i]);
  }
  *key = f.h;
  return PXD_OK;
}

pxd_status pxd_maps_write(const char* dir, const pxd_maps_meta* meta, const int32_t* excluded,
                          const float* maps) {
  if (!dir || !meta || !maps || (meta->n_excluded > 0 && !excluded))
    return err(PXD_ERR_ARG, "maps_write: null argument");
  pxd_status st = check_meta(*meta);
  if (st != PXD_OK) return pxd_set_error_(PXD_ERR_ARG, pxd_last_error());
  for (int i = 1; i < meta->n_excluded; ++i)
    if (excluded[i] <= excluded[i - 1]) return err(PXD_ERR_ARG, "maps_write: excluded atoms must ascend");
  for (int i = 0; i < meta->n_excluded; ++i)
    if (excluded[i] < 0 || excluded[i] >= meta->n_atoms)
      return err(PXD_ERR_ARG, "maps_write: excluded atom %d out of range", excluded[i]);
  const std::string d(dir);
  st = make_dirs(d);
  if (st != PXD_OK) return st;
  // the metadata marks completion: drop an old one before any map changes
  const std::string meta_path = join(d, kMetaName), tmp_path = join(d, kMetaTmp);
  if (unlink(meta_path.c_str()) != 0 && errno != ENOENT)
    return err(PXD_ERR_IO, "maps_write: cannot remove %s: %s", meta_path.c_str(), strerror(errno));
  const long long P = n_points(meta->box);
  std::vector<uint64_t> sum(meta->n_types + 1);
  for (int s = 0; s <= meta->n_types; ++s) {
    const int type = s == 0 ? -1 : meta->types[s - 1];
    st = write_map(join(d, map_file_name(type)), type, meta->box, maps + (size_t)s * P * 4, &sum[s]);
    if (st != PXD_OK) return st;
  }
  FILE* f = fopen(tmp_path.c_str(), "w");
  if (!f) return err(PXD_ERR_IO, "maps_write: cannot write %s: %s", tmp_path.c_str(), strerror(errno));
  const pxd_box& b = meta->box;
  fprintf(f, "pxdock-maps %d\n", PXD_MAPS_VERSION);
  fprintf(f, "init_xyz %.17g %.17g %.17g\n", b.init_xyz[0], b.init_xyz[1], b.init_xyz[2]);
  fprintf(f, "npoints %d %d %d\n", b.npoints[0], b.npoints[1], b.npoints[2]);
  fprintf(f, "spacing %.17g\n", b.spacing);
  fprintf(f, "slope %.17g\n", meta->slope);
  fprintf(f, "receptor %d %016" PRIx64 "\n", meta->n_atoms, meta->receptor_key);
  fprintf(f, "excluded %d", meta->n_excluded);
  for (int i = 0; i < meta->n_excluded; ++i) fprintf(f, " %d", excluded[i]);
  fprintf(f, "\ntypes %d", meta->n_types);
  for (int t = 0; t < meta->n_types; ++t) fprintf(f, " %d", meta->types[t]);
  fprintf(f, "\n");
  for (int s = 0; s <= meta->n_types; ++s) {
    const int type = s == 0 ? -1 : meta->types[s - 1];
    if (type < 0)
      fprintf(f, "map coul %s %016" PRIx64 "\n", map_file_name(type).c_str(), sum[s]);
    else
      fprintf(f, "map %d %s %016" PRIx64 "\n", type, map_file_name(type).c_str(), sum[s]);
  }
  fprintf(f, "end\n");
  bool ok = !ferror(f);
  ok = (fclose(f) == 0) && ok;
  if (!ok) return err(PXD_ERR_IO, "maps_write: short write to %s", tmp_path.c_str());
  if (rename(tmp_path.c_str(), meta_path.c_str()) != 0)
    return err(PXD_ERR_IO, "maps_write: cannot rename %s: %s", tmp_path.c_str(), strerror(errno));
  return PXD_OK;
}

pxd_status pxd_maps_info(const char* dir, pxd_maps_meta* meta, int32_t* excluded, int32_t cap) {
  if (!dir || !meta) return err(PXD_ERR_ARG, "maps_info: null argument");
  Meta M;
  pxd_status st = parse_meta(dir, &M);
  if (st != PXD_OK) return st;
  if (excluded) {
    if (cap < M.m.n_excluded)
      return err(PXD_ERR_ARG, "maps_info: %d excluded atoms > capacity %d", M.m.n_excluded, cap);
    std::copy(M.excluded.begin(), M.excluded.end(), excluded);
  }
  *meta = M.m;
  return PXD_OK;
}

pxd_status pxd_maps_read(const char* dir, const pxd_maps_meta* expect, const int32_t* excluded,
                         const int32_t* types, int32_t n_types, float* maps, size_t n_floats) {
  if (!dir || !maps || (types && n_types < 0) || (expect && expect->n_excluded > 0 && !excluded))
    return err(PXD_ERR_ARG, "maps_read: null argument");
  Meta M;
  pxd_status st = parse_meta(dir, &M);
  if (st != PXD_OK) return st;
  if (expect) {
    if (expect->n_atoms != M.m.n_atoms || expect->receptor_key != M.m.receptor_key)
      return err(PXD_ERR_CONF, "maps: saved for another receptor (%d atoms, key %016" PRIx64
                 ") than this one (%d atoms, key %016" PRIx64 ")", M.m.n_atoms, M.m.receptor_key,
                 expect->n_atoms, expect->receptor_key);
    if (expect->n_excluded != M.m.n_excluded ||
        !std::equal(M.excluded.begin(), M.excluded.end(), excluded))
      return err(PXD_ERR_CONF, "maps: saved with other movable receptor atoms (%d) than this "
                 "receptor's leaves (%d)", M.m.n_excluded, expect->n_excluded);
  }
  std::vector<int> want;  // meta type slots to load, in output order
  if (types) {
    for (int i = 0; i < n_types; ++i) {
      if (types[i] < 0 || types[i] >= PXD_NUM_VDW_TYPES)
        return err(PXD_ERR_ARG, "maps_read: type %d out of range", types[i]);
      int slot = -1;
      for (int s = 0; s < M.m.n_types; ++s)
        if (M.m.types[s] == types[i]) slot = s + 1;
      if (slot < 0) return err(PXD_ERR_IO, "Map file for atom type [%s] is missing!", kTypeTag[types[i]]);
      want.push_back(slot);
    }
  } else {
    for (int s = 1; s <= M.m.n_types; ++s) want.push_back(s);
  }
  const long long P = n_points(M.m.box);
  const size_t need = (size_t)(want.size() + 1) * P * 4;
  if (n_floats < need) return err(PXD_ERR_ARG, "maps_read: buffer %zu < %zu floats", n_floats, need);
  st = read_map(join(dir, map_file_name(-1)), -1, M.m.box, M.sum[0], maps);
  for (size_t i = 0; st == PXD_OK && i < want.size(); ++i) {
    const int type = M.m.types[want[i] - 1];
    st = read_map(join(dir, map_file_name(type)), type, M.m.box, M.sum[want[i]],
                  maps + (i + 1) * (size_t)P * 4);
  }
  return st;
}

pxd_status pxd_grid_save(const pxd_grid* g, const char* dir) {
  if (!g || !dir) return err(PXD_ERR_ARG, "grid_save: null argument");
  const pxd_receptor* r = g->receptor;
  pxd_maps_meta m;
  memset(&m, 0, sizeof(m));
  m.box = g->box;
  m.slope = g->slope;
  m.n_atoms = r->n_atoms;
  m.n_excluded = (int32_t)r->movable.size();
  m.receptor_key = r->key;
  // maps in ascending type order, as the metadata lists them
  std::vector<int> order(g->n_types);
  for (int t = 0; t < g->n_types; ++t) order[t] = t;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return g->types[a] < g->types[b]; });
  m.n_types = g->n_types;
  for (int t = 0; t < g->n_types; ++t) m.types[t] = g->types[order[t]];
  const size_t per = (size_t)g->dev.P * 4;
  std::vector<float> dev((size_t)(g->n_types + 1) * per), sorted(dev.size());
  pxd_status st = pxd_grid_download(g, dev.data(), dev.size());
  if (st != PXD_OK) return st;
  std::copy(dev.begin(), dev.begin() + per, sorted.begin());
  for (int t = 0; t < g->n_types; ++t)
    std::copy(dev.begin() + (order[t] + 1) * per, dev.begin() + (order[t] + 2) * per,
              sorted.begin() + (t + 1) * per);
  return pxd_maps_write(dir, &m, r->movable.data(), sorted.data());
}

pxd_status pxd_grid_load(const pxd_receptor* r, const char* dir, const int32_t* types,
                         int32_t n_types, double slope, pxd_grid** out) {
  if (!r || !dir || !out || (types && n_types < 0)) return err(PXD_ERR_ARG, "grid_load: null argument");
  *out = nullptr;
  pxd_maps_meta saved;
  pxd_status st = pxd_maps_info(dir, &saved, nullptr, 0);
  if (st != PXD_OK) return st;
  pxd_maps_meta expect = saved;
  expect.n_atoms = r->n_atoms;
  expect.receptor_key = r->key;
  expect.n_excluded = (int32_t)r->movable.size();
  std::vector<int32_t> sel;
  if (types) {
    sel.assign(types, types + n_types);
  } else {
    sel.assign(saved.types, saved.types + saved.n_types);
  }
  const size_t n = (sel.size() + 1) * (size_t)n_points(saved.box) * 4;
  std::vector<float> maps(n);
  // a non-null list, also when it is empty (a null list asks for every type)
  const int32_t* want = sel.empty() ? saved.types : sel.data();
  st = pxd_maps_read(dir, &expect, r->movable.data(), want, (int32_t)sel.size(), maps.data(), n);
  if (st != PXD_OK) return st;
  return pxd_grid_upload(r, &saved.box, sel.data(), (int32_t)sel.size(),
                         slope < 0 ? saved.slope : slope, maps.data(), out);
}

}  // extern "C"
