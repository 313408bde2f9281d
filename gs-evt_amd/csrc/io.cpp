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
#include <errno.h>
#include <fcntl.h>
#include <omp.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

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
    if (q + s + cs > end) return -1;
    const uint64_t c = be ? load_be(q + s, cs) : load_le(q + s, cs);
    const double cnt = bits_to_double(c, pr.count_type);
    if (cnt < 0) return -1;
    s += cs + (long long)cnt * type_size(pr.type);
  }
  return (q + s <= end) ? s : -1;
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
    while (s < N && base[s - 1] != '\n') ++s;
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

}  // extern "C"
