// Host-side readers of the tracker's on-disk inputs (C ABI, no device code;
// built with the host compiler into libgsevt_io.so):
//
// * 3DGS PLY maps as GaussianModel.load_ply reads them
//   (gaussian_splatting/scene/gaussian_model.py:238-327): the vertex element's
//   x/y/z, nx/ny/nz, f_dc_0..2, f_rest_* (sorted by index), opacity, scale_*
//   and rot_* (sorted by index), converted to float32.  ascii,
//   binary_little_endian and binary_big_endian files with any property order
//   and scalar type; elements before the vertex element (with list
//   properties) are skipped.
// * events.txt streams as load_events_from_txt reads them
//   (utils/event_camera/event.py:11-39): one "ts x y p" line of integers per
//   event.  The file is memory-mapped and parsed by OpenMP threads over
//   newline-aligned chunks.
// * ROS1 bags (format 2.0) as load_events_from_bag reads them through rosbag
//   (utils/event_camera/event.py:42-69): the EventArray messages of one
//   topic, found through the bag's index; chunks decompressed by OpenMP
//   threads.
#include <dlfcn.h>
#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <omp.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "../../include/gsevt_io.h"

namespace {

thread_local char g_err[512] = "";

void fail(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ------------------------------------------------------------------ mapping
struct Mapped {
  const char* p = nullptr;
  size_t n = 0;
  int fd = -1;
  ~Mapped() {
    if (p && n) munmap((void*)p, n);
    if (fd >= 0) close(fd);
  }
  bool open(const char* path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) {
      fail("cannot open %s: %s", path, strerror(errno));
      return false;
    }
    struct stat st;
    if (fstat(fd, &st) != 0) {
      fail("cannot stat %s: %s", path, strerror(errno));
      return false;
    }
    n = (size_t)st.st_size;
    if (n == 0) return true;
    void* m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) {
      fail("cannot map %s: %s", path, strerror(errno));
      n = 0;
      return false;
    }
    p = (const char*)m;
    return true;
  }
};

// ---------------------------------------------------------------------- PLY
enum PlyType { T_I8, T_U8, T_I16, T_U16, T_I32, T_U32, T_F32, T_F64, T_BAD };

int type_size(int t) {
  static const int s[] = {1, 1, 2, 2, 4, 4, 4, 8};
  return t >= 0 && t < T_BAD ? s[t] : 0;
}

int parse_type(const std::string& s) {
  if (s == "char" || s == "int8") return T_I8;
  if (s == "uchar" || s == "uint8") return T_U8;
  if (s == "short" || s == "int16") return T_I16;
  if (s == "ushort" || s == "uint16") return T_U16;
  if (s == "int" || s == "int32") return T_I32;
  if (s == "uint" || s == "uint32") return T_U32;
  if (s == "float" || s == "float32") return T_F32;
  if (s == "double" || s == "float64") return T_F64;
  return T_BAD;
}

struct PlyProp {
  std::string name;
  int type = T_BAD;
  bool list = false;
  int count_type = T_BAD;
};

struct PlyElem {
  std::string name;
  long long count = 0;
  std::vector<PlyProp> props;
};

struct PlyFile {
  Mapped m;
  int format = -1;   // 0 ascii, 1 binary little endian, 2 binary big endian
  std::vector<PlyElem> elems;
  size_t body = 0;   // offset of the first data byte
};

bool next_line(const char* p, size_t n, size_t& pos, std::string& line) {
  if (pos >= n) return false;
  size_t e = pos;
  while (e < n && p[e] != '\n') ++e;
  line.assign(p + pos, e - pos);
  if (!line.empty() && line.back() == '\r') line.pop_back();
  pos = e < n ? e + 1 : e;
  return true;
}

std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t')) ++i;
    size_t j = i;
    while (j < s.size() && s[j] != ' ' && s[j] != '\t') ++j;
    if (j > i) out.emplace_back(s.substr(i, j - i));
    i = j;
  }
  return out;
}

bool ply_open(const char* path, PlyFile& f) {
  if (!f.m.open(path)) return false;
  const char* p = f.m.p;
  const size_t n = f.m.n;
  size_t pos = 0;
  std::string line;
  if (!next_line(p, n, pos, line) || line != "ply") {
    fail("%s: not a PLY file", path);
    return false;
  }
  bool ended = false;
  while (next_line(p, n, pos, line)) {
    const std::vector<std::string> t = split_ws(line);
    if (t.empty()) continue;
    if (t[0] == "format") {
      if (t.size() < 2) break;
      if (t[1] == "ascii") f.format = 0;
      else if (t[1] == "binary_little_endian") f.format = 1;
      else if (t[1] == "binary_big_endian") f.format = 2;
      else {
        fail("%s: unknown PLY format %s", path, t[1].c_str());
        return false;
      }
    } else if (t[0] == "element") {
      if (t.size() < 3) {
        fail("%s: bad element line", path);
        return false;
      }
      PlyElem e;
      e.name = t[1];
      e.count = atoll(t[2].c_str());
      if (e.count < 0) {
        fail("%s: negative element count", path);
        return false;
      }
      f.elems.push_back(e);
    } else if (t[0] == "property") {
      if (f.elems.empty()) {
        fail("%s: property before any element", path);
        return false;
      }
      PlyProp pr;
      if (t.size() >= 5 && t[1] == "list") {
        pr.list = true;
        pr.count_type = parse_type(t[2]);
        pr.type = parse_type(t[3]);
        pr.name = t[4];
      } else if (t.size() >= 3) {
        pr.type = parse_type(t[1]);
        pr.name = t[2];
      }
      if (pr.type == T_BAD || (pr.list && pr.count_type == T_BAD)) {
        fail("%s: unsupported property line '%s'", path, line.c_str());
        return false;
      }
      f.elems.back().props.push_back(pr);
    } else if (t[0] == "end_header") {
      ended = true;
      break;
    }
  }
  if (!ended || f.format < 0) {
    fail("%s: incomplete PLY header", path);
    return false;
  }
  f.body = pos;
  return true;
}

inline uint64_t load_be(const unsigned char* q, int size) {
  uint64_t v = 0;
  for (int i = 0; i < size; ++i) v = (v << 8) | q[i];
  return v;
}

inline uint64_t load_le(const unsigned char* q, int size) {
  uint64_t v = 0;
  for (int i = size - 1; i >= 0; --i) v = (v << 8) | q[i];
  return v;
}

double bits_to_double(uint64_t v, int type) {
  switch (type) {
    case T_I8: return (double)(int8_t)v;
    case T_U8: return (double)(uint8_t)v;
    case T_I16: return (double)(int16_t)v;
    case T_U16: return (double)(uint16_t)v;
    case T_I32: return (double)(int32_t)v;
    case T_U32: return (double)(uint32_t)v;
    case T_F32: {
      uint32_t u = (uint32_t)v;
      float x;
      memcpy(&x, &u, 4);
      return (double)x;
    }
    default: {
      double x;
      memcpy(&x, &v, 8);
      return x;
    }
  }
}

// Value of a scalar property read in its own type, then rounded once to
// float32 (float32 stays exact; float64 rounds to nearest, as numpy does).
inline float to_f32(uint64_t v, int type) {
  if (type == T_F32) {
    uint32_t u = (uint32_t)v;
    float x;
    memcpy(&x, &u, 4);
    return x;
  }
  return (float)bits_to_double(v, type);
}

// Size in bytes of one binary record of element e starting at q (list
// properties make records variable), or -1 past the end.
long long binary_record_size(const PlyElem& e, const unsigned char* q, const unsigned char* end, bool be) {
  long long s = 0;
  for (const PlyProp& pr : e.props) {
    if (!pr.list) {
      s += type_size(pr.type);
      continue;
    }
    const int cs = type_size(pr.count_type);
    const long long avail = (long long)(end - q);
    if (s + cs > avail) return -1;
    const uint64_t c = be ? load_be(q + s, cs) : load_le(q + s, cs);
    const double cnt = bits_to_double(c, pr.count_type);
    // a count must be a finite non-negative integer (float and double count
    // types can hold NaN, inf or fractions), and its items must fit in what
    // is left of the file: bounding it first keeps the products below small
    if (!(cnt >= 0) || cnt != floor(cnt)) return -1;
    const long long ts = type_size(pr.type);
    if (ts <= 0 || cnt > (double)((avail - s - cs) / ts)) return -1;
    s += cs + (long long)cnt * ts;
  }
  return s <= (long long)(end - q) ? s : -1;
}

// Targets of the 3DGS vertex properties (index into the output arrays).
struct VertexMap {
  int vi = -1;                    // index of the vertex element
  std::vector<int> role;          // per vertex property: output slot, -1 unused
  int n_rest = 0, n_scale = 0, n_rot = 0;
};

enum Slot { S_X = 0, S_Y, S_Z, S_NX, S_NY, S_NZ, S_DC0, S_DC1, S_DC2, S_OP, S_REST = 100, S_SCALE = 10000, S_ROT = 20000 };

bool suffix_index(const std::string& name, const char* prefix, int& idx) {
  const size_t L = strlen(prefix);
  if (name.compare(0, L, prefix) != 0) return false;
  // the reference sorts by int(name.split("_")[-1])
  const size_t us = name.rfind('_');
  if (us == std::string::npos || us + 1 >= name.size()) return false;
  char* endp = nullptr;
  const long v = strtol(name.c_str() + us + 1, &endp, 10);
  if (*endp != '\0' || v < 0 || v > 9999) return false;
  idx = (int)v;
  return true;
}

bool build_vertex_map(const char* path, const PlyFile& f, VertexMap& vm) {
  for (size_t i = 0; i < f.elems.size(); ++i)
    if (f.elems[i].name == "vertex") { vm.vi = (int)i; break; }
  if (vm.vi < 0) {
    fail("%s: no vertex element", path);
    return false;
  }
  const PlyElem& e = f.elems[vm.vi];
  std::vector<int> rest, scale, rot;
  vm.role.assign(e.props.size(), -1);
  bool seen[10] = {false};
  static const char* fixed[10] = {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2", "opacity"};
  for (size_t k = 0; k < e.props.size(); ++k) {
    const PlyProp& pr = e.props[k];
    for (int s = 0; s < 10; ++s)
      if (pr.name == fixed[s]) {
        if (pr.list) {
          fail("%s: property %s is a list", path, pr.name.c_str());
          return false;
        }
        vm.role[k] = s;
        seen[s] = true;
      }
    int idx;
    if (suffix_index(pr.name, "f_rest_", idx)) { vm.role[k] = S_REST + idx; rest.push_back(idx); }
    else if (suffix_index(pr.name, "scale_", idx)) { vm.role[k] = S_SCALE + idx; scale.push_back(idx); }
    else if (pr.name.compare(0, 3, "rot") == 0 && suffix_index(pr.name, "rot", idx)) {
      vm.role[k] = S_ROT + idx;
      rot.push_back(idx);
    }
    if (vm.role[k] >= S_REST && pr.list) {
      fail("%s: property %s is a list", path, pr.name.c_str());
      return false;
    }
  }
  static const int needed[] = {S_X, S_Y, S_Z, S_DC0, S_DC1, S_DC2, S_OP};
  for (int s : needed)
    if (!seen[s]) {
      fail("%s: vertex property %s missing", path, fixed[s]);
      return false;
    }
  // indices must be 0..n-1 (the reference fills columns in sorted order)
  auto dense = [&](std::vector<int>& v, const char* what) {
    std::vector<char> hit(v.size(), 0);
    for (int i : v) {
      if (i >= (int)v.size() || hit[i]) {
        fail("%s: %s properties are not numbered 0..%zu", path, what, v.size() - 1);
        return false;
      }
      hit[i] = 1;
    }
    return true;
  };
  if (!dense(rest, "f_rest_") || !dense(scale, "scale_") || !dense(rot, "rot_")) return false;
  vm.n_rest = (int)rest.size();
  vm.n_scale = (int)scale.size();
  vm.n_rot = (int)rot.size();
  return true;
}

struct Outs {
  float *xyz, *normals, *f_dc, *f_rest, *opacity, *scales, *rots;
};

inline void store(const VertexMap& vm, const Outs& o, long long v, int role, float x) {
  if (role < 0) return;
  if (role < 3) o.xyz[3 * v + role] = x;
  else if (role < 6) { if (o.normals) o.normals[3 * v + role - 3] = x; }
  else if (role < 9) o.f_dc[3 * v + role - 6] = x;
  else if (role == S_OP) o.opacity[v] = x;
  else if (role >= S_ROT) o.rots[(size_t)v * vm.n_rot + role - S_ROT] = x;
  else if (role >= S_SCALE) o.scales[(size_t)v * vm.n_scale + role - S_SCALE] = x;
  else o.f_rest[(size_t)v * vm.n_rest + role - S_REST] = x;
}

bool read_vertices_binary(const char* path, const PlyFile& f, const VertexMap& vm, const Outs& o) {
  const bool be = f.format == 2;
  const unsigned char* q = (const unsigned char*)f.m.p + f.body;
  const unsigned char* end = (const unsigned char*)f.m.p + f.m.n;
  for (int ei = 0; ei < vm.vi; ++ei) {   // skip earlier elements
    const PlyElem& e = f.elems[ei];
    for (long long r = 0; r < e.count; ++r) {
      const long long s = binary_record_size(e, q, end, be);
      if (s < 0) {
        fail("%s: truncated element %s", path, e.name.c_str());
        return false;
      }
      q += s;
    }
  }
  const PlyElem& e = f.elems[vm.vi];
  bool fixed_size = true;
  long long stride = 0;
  std::vector<int> off(e.props.size());
  for (size_t k = 0; k < e.props.size(); ++k) {
    if (e.props[k].list) fixed_size = false;
    off[k] = (int)stride;
    stride += type_size(e.props[k].type);
  }
  if (fixed_size) {
    if (e.count > 0 && (long long)(end - q) / stride < e.count) {
      fail("%s: vertex data truncated (%lld records of %lld bytes expected)", path, e.count, stride);
      return false;
    }
#pragma omp parallel for schedule(static)
    for (long long v = 0; v < e.count; ++v) {
      const unsigned char* r = q + v * stride;
      for (size_t k = 0; k < e.props.size(); ++k) {
        if (vm.role[k] < 0) continue;
        const int ts = type_size(e.props[k].type);
        const uint64_t bits = be ? load_be(r + off[k], ts) : load_le(r + off[k], ts);
        store(vm, o, v, vm.role[k], to_f32(bits, e.props[k].type));
      }
    }
    return true;
  }
  for (long long v = 0; v < e.count; ++v) {   // vertex records with lists: sequential walk
    const long long s = binary_record_size(e, q, end, be);
    if (s < 0) {
      fail("%s: vertex data truncated at record %lld", path, v);
      return false;
    }
    long long at = 0;
    for (size_t k = 0; k < e.props.size(); ++k) {
      const PlyProp& pr = e.props[k];
      if (pr.list) {
        const int cs = type_size(pr.count_type);
        const uint64_t c = be ? load_be(q + at, cs) : load_le(q + at, cs);
        at += cs + (long long)bits_to_double(c, pr.count_type) * type_size(pr.type);
        continue;
      }
      const int ts = type_size(pr.type);
      const uint64_t bits = be ? load_be(q + at, ts) : load_le(q + at, ts);
      store(vm, o, v, vm.role[k], to_f32(bits, pr.type));
      at += ts;
    }
    q += s;
  }
  return true;
}

bool read_vertices_ascii(const char* path, const PlyFile& f, const VertexMap& vm, const Outs& o) {
  const char* p = f.m.p;
  const size_t n = f.m.n;
  size_t pos = f.body;
  std::string line;
  for (int ei = 0; ei < vm.vi; ++ei)
    for (long long r = 0; r < f.elems[ei].count; ++r)
      if (!next_line(p, n, pos, line)) {
        fail("%s: truncated element %s", path, f.elems[ei].name.c_str());
        return false;
      }
  const PlyElem& e = f.elems[vm.vi];
  for (long long v = 0; v < e.count; ++v) {
    if (!next_line(p, n, pos, line)) {
      fail("%s: vertex data truncated at record %lld", path, v);
      return false;
    }
    const std::vector<std::string> t = split_ws(line);
    size_t at = 0;
    for (size_t k = 0; k < e.props.size(); ++k) {
      const PlyProp& pr = e.props[k];
      if (pr.list) {
        if (at >= t.size()) break;
        at += 1 + (size_t)atoll(t[at].c_str());
        continue;
      }
      if (at >= t.size()) {
        fail("%s: vertex record %lld has too few values", path, v);
        return false;
      }
      // parsed in the property's own type, then rounded once to float32
      const double x = strtod(t[at].c_str(), nullptr);
      const float xf = (pr.type == T_F32) ? strtof(t[at].c_str(), nullptr) : (float)x;
      store(vm, o, v, vm.role[k], xf);
      ++at;
    }
  }
  return true;
}

// ------------------------------------------------------------------- events
// Parses one non-negative or negative decimal integer at p (no leading
// blanks); returns false when there is none.
inline bool parse_int(const char*& p, const char* end, long long& v) {
  bool neg = false;
  if (p < end && (*p == '-' || *p == '+')) {
    neg = *p == '-';
    ++p;
  }
  if (p >= end || *p < '0' || *p > '9') return false;
  long long x = 0;
  while (p < end && *p >= '0' && *p <= '9') {
    x = x * 10 + (*p - '0');
    ++p;
  }
  v = neg ? -x : x;
  return true;
}

inline bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// One event line as the reference reads it: [int(t) for t in
// line.strip().split(' ')], then fields 0..3.  Returns 0 (ok), 1 (a field is
// not an integer: ValueError there) or 2 (fewer than 4 fields: IndexError).
// Extra integer fields are ignored, as there.
inline int parse_event_line(const char* p, const char* e, long long* out4) {
  while (e > p && is_ws(e[-1])) --e;
  while (p < e && is_ws(*p)) ++p;
  int k = 0;
  for (;;) {
    const char* te = (const char*)memchr(p, ' ', (size_t)(e - p));
    if (!te) te = e;
    const char* a = p;
    const char* b = te;
    while (a < b && is_ws(*a)) ++a;   // int() accepts blanks around the digits
    while (b > a && is_ws(b[-1])) --b;
    long long v;
    const char* q = a;
    if (!parse_int(q, b, v) || q != b) return 1;
    if (k < 4) out4[k] = v;
    ++k;
    if (te == e) break;
    p = te + 1;
  }
  return k < 4 ? 2 : 0;
}

// ------------------------------------------------------------------ ROS bag
// Format 2.0 (the rosbag file format): "#ROSBAG V2.0\n", then records of
//   u32 header_len | header fields (u32 len, "name=" + value bytes) | u32 data_len | data
// with the header field op: 0x02 message data, 0x03 bag header, 0x04 index
// data, 0x05 chunk, 0x06 chunk info, 0x07 connection.  Integers are
// little-endian; a time is u32 secs then u32 nsecs.

inline uint16_t rd16(const unsigned char* q) { return (uint16_t)(q[0] | (q[1] << 8)); }
inline uint32_t rd32(const unsigned char* q) {
  return (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
}
inline uint64_t rd64(const unsigned char* q) { return (uint64_t)rd32(q) | ((uint64_t)rd32(q + 4) << 32); }

struct BagRecord {
  const unsigned char* hdr = nullptr;
  uint32_t hdr_len = 0;
  const unsigned char* data = nullptr;
  uint32_t data_len = 0;
  size_t next = 0;   // offset just past the record
};

// The record at offset pos of buf[0, n); false when it runs past the end.
bool bag_record(const unsigned char* buf, size_t n, size_t pos, BagRecord& r) {
  if (pos > n || n - pos < 4) return false;
  r.hdr_len = rd32(buf + pos);
  size_t q = pos + 4;
  if (r.hdr_len > n - q) return false;
  r.hdr = buf + q;
  q += r.hdr_len;
  if (n - q < 4) return false;
  r.data_len = rd32(buf + q);
  q += 4;
  if (r.data_len > n - q) return false;
  r.data = buf + q;
  r.next = q + r.data_len;
  return true;
}

// Value of the field `name` in a run of (u32 len, "name=value") fields, or
// nullptr; *len receives its size.  Record headers and connection headers
// share this encoding.
const unsigned char* bag_field(const unsigned char* h, uint32_t n, const char* name, uint32_t* len) {
  const size_t nl = strlen(name);
  uint32_t q = 0;
  while (n - q >= 4) {
    const uint32_t fl = rd32(h + q);
    q += 4;
    if (fl > n - q) return nullptr;
    const unsigned char* f = h + q;
    const unsigned char* eq = (const unsigned char*)memchr(f, '=', fl);
    if (eq && (size_t)(eq - f) == nl && memcmp(f, name, nl) == 0) {
      *len = (uint32_t)(f + fl - eq - 1);
      return eq + 1;
    }
    q += fl;
  }
  return nullptr;
}

bool bag_u32(const BagRecord& r, const char* name, uint32_t& v) {
  uint32_t len = 0;
  const unsigned char* p = bag_field(r.hdr, r.hdr_len, name, &len);
  if (!p || len != 4) return false;
  v = rd32(p);
  return true;
}

bool bag_u64(const BagRecord& r, const char* name, uint64_t& v) {
  uint32_t len = 0;
  const unsigned char* p = bag_field(r.hdr, r.hdr_len, name, &len);
  if (!p || len != 8) return false;
  v = rd64(p);
  return true;
}

int bag_op(const BagRecord& r) {
  uint32_t len = 0;
  const unsigned char* p = bag_field(r.hdr, r.hdr_len, "op", &len);
  return p && len == 1 ? p[0] : -1;
}

std::string field_str(const unsigned char* h, uint32_t n, const char* name, bool* found = nullptr) {
  uint32_t len = 0;
  const unsigned char* p = bag_field(h, n, name, &len);
  if (found) *found = p != nullptr;
  return p ? std::string((const char*)p, len) : std::string();
}

// Field lines ("type", "name") of every section of a message definition: the
// top message under "", the embedded ones under the name after "MSG: ".
// Comments, blank lines, constants and the "===" separators are dropped.
std::vector<std::pair<std::string, std::vector<std::pair<std::string, std::string>>>> def_sections(
    const std::string& def) {
  std::vector<std::pair<std::string, std::vector<std::pair<std::string, std::string>>>> secs(1);
  size_t pos = 0;
  while (pos <= def.size()) {
    size_t e = def.find('\n', pos);
    if (e == std::string::npos) e = def.size();
    std::string line = def.substr(pos, e - pos);
    pos = e + 1;
    if (line.compare(0, 4, "MSG:") == 0) {
      std::vector<std::string> w = split_ws(line.substr(4));
      secs.push_back({w.empty() ? std::string() : w[0], {}});
      continue;
    }
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    if (line.find('=') != std::string::npos) continue;   // constant or separator
    std::vector<std::string> w = split_ws(line);
    if (w.size() >= 2) secs.back().second.push_back({w[0], w[1]});
  }
  return secs;
}

// Whether a connection's messages have the wire layout of dvs_msgs/EventArray
// (= prophesee_event_msgs/EventArray):
//   Header header {u32 seq; time stamp; string frame_id}; u32 height; u32 width;
//   Event[] events {u16 x; u16 y; time ts; bool polarity}.
// Read from the stored message definition; a connection without one is
// accepted by type name.
bool event_array_layout(const std::string& type, const std::string& def, std::string& why) {
  if (def.empty()) {
    if (type == "dvs_msgs/EventArray" || type == "prophesee_event_msgs/EventArray") return true;
    why = "type " + type + " has no message definition";
    return false;
  }
  typedef std::vector<std::pair<std::string, std::string>> Fields;
  const auto secs = def_sections(def);
  auto find = [&](const std::string& name) -> const Fields* {
    for (size_t i = 1; i < secs.size(); ++i)
      if (secs[i].first == name) return &secs[i].second;
    return nullptr;
  };
  const std::string pkg = type.substr(0, type.find('/'));
  const Fields& top = secs[0].second;
  if (top.size() != 4 || top[0].second != "header" ||
      (top[0].first != "Header" && top[0].first != "std_msgs/Header") ||
      top[1] != Fields::value_type("uint32", "height") || top[2] != Fields::value_type("uint32", "width") ||
      top[3].second != "events" || top[3].first.size() < 3 ||
      top[3].first.compare(top[3].first.size() - 2, 2, "[]") != 0) {
    why = "type " + type + " is not {Header header, uint32 height, uint32 width, Event[] events}";
    return false;
  }
  std::string ev = top[3].first.substr(0, top[3].first.size() - 2);
  if (ev.find('/') == std::string::npos) ev = pkg + "/" + ev;
  const Fields* hf = find("std_msgs/Header");
  const Fields* ef = find(ev);
  const Fields want_h = {{"uint32", "seq"}, {"time", "stamp"}, {"string", "frame_id"}};
  const Fields want_e = {{"uint16", "x"}, {"uint16", "y"}, {"time", "ts"}, {"bool", "polarity"}};
  if (!hf || *hf != want_h) {
    why = "type " + type + ": std_msgs/Header is not {uint32 seq, time stamp, string frame_id}";
    return false;
  }
  if (!ef || *ef != want_e) {
    why = "type " + type + ": " + ev + " is not {uint16 x, uint16 y, time ts, bool polarity}";
    return false;
  }
  return true;
}

// bz2 and lz4 chunks go through the system libbz2 / liblz4, loaded on first
// use by their documented C entry points (BZ2_bzBuffToBuffDecompress; the
// LZ4F frame decoder, which reads the frames rosbag's lz4 chunks hold).
typedef int (*Bz2Decompress)(char*, unsigned int*, char*, unsigned int, int, int);
typedef size_t (*Lz4fCreate)(void**, unsigned);
typedef size_t (*Lz4fDecompress)(void*, void*, size_t*, const void*, size_t*, const void*);
typedef size_t (*Lz4fFree)(void*);
typedef unsigned (*Lz4fIsError)(size_t);
typedef const char* (*Lz4fErrorName)(size_t);

void* open_lib(const char* a, const char* b) {
  void* h = dlopen(a, RTLD_NOW | RTLD_LOCAL);
  return h ? h : dlopen(b, RTLD_NOW | RTLD_LOCAL);
}

Bz2Decompress bz2_decompress() {
  static const Bz2Decompress f = [] {
    void* h = open_lib("libbz2.so.1", "libbz2.so");
    return h ? (Bz2Decompress)dlsym(h, "BZ2_bzBuffToBuffDecompress") : (Bz2Decompress) nullptr;
  }();
  return f;
}

struct Lz4f {
  Lz4fCreate create = nullptr;
  Lz4fDecompress decompress = nullptr;
  Lz4fFree free_ctx = nullptr;
  Lz4fIsError is_error = nullptr;
  Lz4fErrorName error_name = nullptr;
  bool ok() const { return create && decompress && free_ctx && is_error && error_name; }
};

const Lz4f& lz4f() {
  static const Lz4f f = [] {
    Lz4f r;
    void* h = open_lib("liblz4.so.1", "liblz4.so");
    if (h) {
      r.create = (Lz4fCreate)dlsym(h, "LZ4F_createDecompressionContext");
      r.decompress = (Lz4fDecompress)dlsym(h, "LZ4F_decompress");
      r.free_ctx = (Lz4fFree)dlsym(h, "LZ4F_freeDecompressionContext");
      r.is_error = (Lz4fIsError)dlsym(h, "LZ4F_isError");
      r.error_name = (Lz4fErrorName)dlsym(h, "LZ4F_getErrorName");
    }
    return r;
  }();
  return f;
}

// Uncompressed bytes of a chunk record's data into out (size bytes exactly).
bool chunk_data(const std::string& comp, const unsigned char* src, uint32_t n, uint32_t size,
                std::vector<unsigned char>& out, std::string& err) {
  out.resize(size);
  if (comp == "bz2") {
    Bz2Decompress f = bz2_decompress();
    if (!f) {
      err = "bz2 chunk, but libbz2.so.1 cannot be loaded";
      return false;
    }
    unsigned int got = size;
    const int rc = f((char*)out.data(), &got, (char*)src, n, 0, 0);
    if (rc != 0 || got != size) {
      err = "bz2 chunk does not decompress to its size (code " + std::to_string(rc) + ")";
      return false;
    }
    return true;
  }
  if (comp == "lz4") {
    const Lz4f& z = lz4f();
    if (!z.ok()) {
      err = "lz4 chunk, but liblz4.so.1 cannot be loaded";
      return false;
    }
    void* ctx = nullptr;
    if (z.is_error(z.create(&ctx, 100u))) {
      err = "lz4 context";
      return false;
    }
    size_t in = 0, outp = 0, hint = 1;
    while (hint != 0 && in < n) {
      size_t dn = size - outp, sn = n - in;
      hint = z.decompress(ctx, out.data() + outp, &dn, src + in, &sn, nullptr);
      if (z.is_error(hint)) {
        err = std::string("lz4 chunk: ") + z.error_name(hint);
        z.free_ctx(ctx);
        return false;
      }
      in += sn;
      outp += dn;
      if (sn == 0 && dn == 0) break;
    }
    z.free_ctx(ctx);
    if (hint != 0 || outp != size) {
      err = "lz4 chunk does not decompress to its size";
      return false;
    }
    return true;
  }
  err = "chunk compression '" + comp + "' is not none, bz2 or lz4";
  return false;
}

// A kept message; the list is built in file order (chunks by position, records
// in chunk order), so a stable sort by bag time leaves ties in file order.
struct BagMsg {
  uint32_t secs, nsecs;          // bag (record) time
  const unsigned char* events;   // n_events records of 13 bytes
  uint32_t n_events;
  uint32_t stamp_secs, stamp_nsecs;   // header.stamp
};

struct BagScan {
  Mapped m;
  std::vector<std::vector<unsigned char>> bufs;   // uncompressed chunks
  std::vector<BagMsg> msgs;                        // kept messages in bag time order
};

// Messages of `topic` as rosbag's read_messages(topics=[topic]) yields them:
// through the bag's index (an unindexed bag is an error, as there), in bag
// time order, ties in file order; messages whose bag time Time.to_sec() =
// secs + nsecs / 1e9 (in double) is below start_time dropped when has_start.
bool bag_scan(const char* path, const char* topic, int has_start, double start_time, BagScan& s) {
  if (!s.m.open(path)) return false;
  const unsigned char* b = (const unsigned char*)s.m.p;
  const size_t n = s.m.n;
  static const char kMagic[] = "#ROSBAG V2.0\n";
  const size_t ml = sizeof(kMagic) - 1;
  if (n < ml || memcmp(b, kMagic, ml) != 0) {
    fail("%s: not a ROS bag of format 2.0", path);
    return false;
  }
  BagRecord r;
  uint64_t index_pos = 0;
  if (!bag_record(b, n, ml, r) || bag_op(r) != 0x03 || !bag_u64(r, "index_pos", index_pos)) {
    fail("%s: no bag header record", path);
    return false;
  }
  if (index_pos == 0) {
    fail("%s: unindexed bag (its index was never written; rosbag reindex rebuilds it)", path);
    return false;
  }
  if (index_pos > n) {
    fail("%s: index position %llu past the end of the file", path, (unsigned long long)index_pos);
    return false;
  }
  // index section: connection and chunk info records
  std::vector<uint32_t> wanted;
  std::vector<std::pair<uint64_t, std::vector<uint32_t>>> chunk_conns;   // (chunk_pos, connections with messages)
  for (size_t pos = index_pos; pos < n; pos = r.next) {
    if (!bag_record(b, n, pos, r)) {
      fail("%s: truncated index record at offset %zu", path, pos);
      return false;
    }
    const int op = bag_op(r);
    if (op == 0x07) {
      uint32_t conn = 0;
      bool has_topic = false;
      const std::string t = field_str(r.hdr, r.hdr_len, "topic", &has_topic);
      if (!bag_u32(r, "conn", conn) || !has_topic) {
        fail("%s: connection record without conn/topic at offset %zu", path, pos);
        return false;
      }
      if (t != topic) continue;
      const std::string type = field_str(r.data, r.data_len, "type");
      const std::string def = field_str(r.data, r.data_len, "message_definition");
      std::string why;
      if (!event_array_layout(type, def, why)) {
        fail("%s: topic %s: %s", path, topic, why.c_str());
        return false;
      }
      wanted.push_back(conn);
    } else if (op == 0x06) {
      uint64_t cpos = 0;
      uint32_t count = 0;
      if (!bag_u64(r, "chunk_pos", cpos) || !bag_u32(r, "count", count) || (uint64_t)count * 8 > r.data_len) {
        fail("%s: malformed chunk info record at offset %zu", path, pos);
        return false;
      }
      std::vector<uint32_t> conns;
      for (uint32_t i = 0; i < count; ++i)
        if (rd32(r.data + 8 * i + 4) > 0) conns.push_back(rd32(r.data + 8 * i));
      chunk_conns.push_back({cpos, conns});
    }
  }
  auto is_wanted = [&](uint32_t c) {
    for (uint32_t w : wanted)
      if (w == c) return true;
    return false;
  };
  std::vector<uint64_t> chunks;
  for (const auto& cc : chunk_conns)
    for (uint32_t c : cc.second)
      if (is_wanted(c)) {
        chunks.push_back(cc.first);
        break;
      }
  std::sort(chunks.begin(), chunks.end());
  chunks.erase(std::unique(chunks.begin(), chunks.end()), chunks.end());
  const long long nc = (long long)chunks.size();
  s.bufs.assign(nc, {});
  std::vector<std::vector<BagMsg>> per(nc);
  std::vector<std::string> errs(nc);
#pragma omp parallel for schedule(dynamic, 1)
  for (long long c = 0; c < nc; ++c) {
    BagRecord cr;
    uint32_t size = 0;
    if (!bag_record(b, n, chunks[c], cr) || bag_op(cr) != 0x05 || !bag_u32(cr, "size", size)) {
      errs[c] = "no chunk record at offset " + std::to_string(chunks[c]);
      continue;
    }
    const std::string comp = field_str(cr.hdr, cr.hdr_len, "compression");
    const unsigned char* cd = cr.data;
    if (comp != "none") {
      if (!chunk_data(comp, cr.data, cr.data_len, size, s.bufs[c], errs[c])) continue;
      cd = s.bufs[c].data();
    } else if (cr.data_len != size) {
      errs[c] = "uncompressed chunk of " + std::to_string(cr.data_len) + " bytes, size field " + std::to_string(size);
      continue;
    }
    BagRecord mr;
    for (size_t q = 0; q < size; q = mr.next) {
      if (!bag_record(cd, size, q, mr)) {
        errs[c] = "truncated record at offset " + std::to_string(q) + " of the chunk at " + std::to_string(chunks[c]);
        break;
      }
      uint32_t conn = 0;
      uint64_t t = 0;
      if (bag_op(mr) != 0x02 || !bag_u32(mr, "conn", conn) || !is_wanted(conn)) continue;
      if (!bag_u64(mr, "time", t)) {
        errs[c] = "message record without time in the chunk at " + std::to_string(chunks[c]);
        break;
      }
      // header {u32 seq, time stamp, string frame_id}, u32 height, u32 width, u32 n, n x 13 bytes
      const unsigned char* d = mr.data;
      const uint32_t L = mr.data_len;
      const uint32_t f = L >= 16 ? rd32(d + 12) : 0;
      bool ok = L >= 16 && f <= L - 16 && L - 16 - f >= 12;
      uint32_t ne = 0;
      if (ok) {
        ne = rd32(d + 16 + f + 8);
        ok = ne <= (L - 28 - f) / 13;
      }
      if (!ok) {
        errs[c] = "message of connection " + std::to_string(conn) + " at offset " + std::to_string(q) +
                  " of the chunk at " + std::to_string(chunks[c]) + " is shorter than its EventArray";
        break;
      }
      per[c].push_back(BagMsg{(uint32_t)t, (uint32_t)(t >> 32), d + 28 + f, ne, rd32(d + 4), rd32(d + 8)});
    }
  }
  for (long long c = 0; c < nc; ++c)
    if (!errs[c].empty()) {
      fail("%s: %s", path, errs[c].c_str());
      return false;
    }
  for (auto& v : per)
    for (const BagMsg& g : v)
      if (!has_start || !((double)g.secs + (double)g.nsecs / 1e9 < start_time)) s.msgs.push_back(g);
  std::stable_sort(s.msgs.begin(), s.msgs.end(), [](const BagMsg& m1, const BagMsg& m2) {
    return m1.secs != m2.secs ? m1.secs < m2.secs : m1.nsecs < m2.nsecs;
  });
  return true;
}

}  // namespace

extern "C" {

const char* gsevt_io_last_error(void) { return g_err; }

int gsevt_ply_info(const char* path, long long* n_vertices, int* n_rest, int* n_scale, int* n_rot) {
  PlyFile f;
  if (!ply_open(path, f)) return -1;
  VertexMap vm;
  if (!build_vertex_map(path, f, vm)) return -1;
  *n_vertices = f.elems[vm.vi].count;
  *n_rest = vm.n_rest;
  *n_scale = vm.n_scale;
  *n_rot = vm.n_rot;
  return 0;
}

int gsevt_ply_read(const char* path, float* xyz, float* normals, float* f_dc, float* f_rest, float* opacity,
                   float* scales, float* rots) {
  PlyFile f;
  if (!ply_open(path, f)) return -1;
  VertexMap vm;
  if (!build_vertex_map(path, f, vm)) return -1;
  if (!xyz || !f_dc || !opacity || (vm.n_rest && !f_rest) || (vm.n_scale && !scales) || (vm.n_rot && !rots)) {
    fail("gsevt_ply_read: null output array");
    return -1;
  }
  const Outs o{xyz, normals, f_dc, f_rest, opacity, scales, rots};
  const bool ok = f.format == 0 ? read_vertices_ascii(path, f, vm, o) : read_vertices_binary(path, f, vm, o);
  return ok ? 0 : -1;
}

long long gsevt_events_txt_lines(const char* path) {
  Mapped m;
  if (!m.open(path)) return -1;
  long long c = 0;
#pragma omp parallel for reduction(+ : c) schedule(static)
  for (long long i = 0; i < (long long)m.n; ++i) c += m.p[i] == '\n';
  if (m.n > 0 && m.p[m.n - 1] != '\n') ++c;   // last line without a newline
  return c;
}

int gsevt_events_txt_read(const char* path, long long capacity, long long* ts, int* x, int* y, int* p,
                          long long* n_out, long long* bad_line, int* bad_kind) {
  Mapped m;
  if (!m.open(path)) return -1;
  *n_out = 0;
  *bad_line = -1;
  *bad_kind = 0;
  if (m.n == 0) return 0;
  const char* base = m.p;
  const long long N = (long long)m.n;
  const int T = omp_get_max_threads();
  // newline-aligned chunks: chunk t starts after the first '\n' at or past t*N/T
  std::vector<long long> start(T + 1, N);
  start[0] = 0;
  for (int t = 1; t < T; ++t) {
    long long s = (long long)((double)N * t / T);
    if (s < start[t - 1]) s = start[t - 1];
    // s == 0 (a file shorter than the thread count) is a line start already
    while (s > 0 && s < N && base[s - 1] != '\n') ++s;
    start[t] = s;
  }
  std::vector<long long> lines(T, 0), first_bad(T, -1);
  std::vector<int> first_kind(T, 0);
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    long long c = 0;
    for (long long i = start[t]; i < start[t + 1]; ++i) c += base[i] == '\n';
    if (t == T - 1 || start[t + 1] == N) {
      if (start[t + 1] == N && start[t] < N && base[N - 1] != '\n') ++c;
    }
    lines[t] = c;
  }
  std::vector<long long> first(T + 1, 0);
  for (int t = 0; t < T; ++t) first[t + 1] = first[t] + lines[t];
  if (first[T] > capacity) {
    fail("gsevt_events_txt_read: %lld lines exceed the capacity %lld", first[T], capacity);
    return -1;
  }
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    long long li = first[t];
    const char* q = base + start[t];
    const char* qe = base + start[t + 1];
    while (q < qe) {
      const char* e = (const char*)memchr(q, '\n', (size_t)(qe - q));
      const char* le = e ? e : qe;
      long long v[4] = {0, 0, 0, 0};
      int kind = parse_event_line(q, le, v);
      if (kind == 0 && (v[1] < INT32_MIN || v[1] > INT32_MAX || v[2] < INT32_MIN || v[2] > INT32_MAX ||
                        v[3] < INT32_MIN || v[3] > INT32_MAX))
        kind = 3;   // does not fit the int32 columns
      if (kind != 0) {
        if (first_bad[t] < 0) {
          first_bad[t] = li;
          first_kind[t] = kind;
        }
        v[0] = v[1] = v[2] = v[3] = 0;
      }
      ts[li] = v[0];
      x[li] = (int)v[1];
      y[li] = (int)v[2];
      p[li] = (int)v[3];
      ++li;
      q = e ? e + 1 : qe;
    }
  }
  for (int t = 0; t < T; ++t)
    if (first_bad[t] >= 0) {
      *bad_line = first_bad[t];
      *bad_kind = first_kind[t];
      break;
    }
  *n_out = first[T];
  return 0;
}

struct gsevt_bag {
  BagScan scan;
};

gsevt_bag* gsevt_bag_open(const char* path, const char* topic, int has_start, double start_time,
                          long long* n_messages, long long* n_events) {
  gsevt_bag* h = new (std::nothrow) gsevt_bag;
  if (!h) {
    fail("gsevt_bag_open: out of memory");
    return nullptr;
  }
  if (!bag_scan(path, topic, has_start, start_time, h->scan)) {
    delete h;
    return nullptr;
  }
  long long e = 0;
  for (const BagMsg& g : h->scan.msgs) e += g.n_events;
  *n_messages = (long long)h->scan.msgs.size();
  *n_events = e;
  return h;
}

int gsevt_bag_events(const gsevt_bag* h, int event_stamps, long long* msg_events, long long* ts_us, int* x, int* y,
                     int* p) {
  if (!h) {
    fail("gsevt_bag_events: null handle");
    return -1;
  }
  const std::vector<BagMsg>& msgs = h->scan.msgs;
  const long long M = (long long)msgs.size();
  std::vector<long long> first(M + 1, 0);
  for (long long i = 0; i < M; ++i) first[i + 1] = first[i] + msgs[i].n_events;
#pragma omp parallel for schedule(dynamic, 64)
  for (long long i = 0; i < M; ++i) {
    const BagMsg& g = msgs[i];
    msg_events[i] = g.n_events;
    const long long hs = (long long)g.stamp_secs * 1000000ll + g.stamp_nsecs / 1000u;
    long long o = first[i];
    for (uint32_t k = 0; k < g.n_events; ++k, ++o) {
      const unsigned char* q = g.events + 13ull * k;
      x[o] = rd16(q);
      y[o] = rd16(q + 2);
      ts_us[o] = event_stamps ? (long long)rd32(q + 4) * 1000000ll + rd32(q + 8) / 1000u : hs;
      p[o] = q[12] != 0;
    }
  }
  return 0;
}

void gsevt_bag_close(gsevt_bag* h) { delete h; }

}  // extern "C"
