// t3d_gltf.cpp — glTF 2.0 meshes and skins into the resident mesh pool, as
// LoadGLTFData builds them (gltf.go:560-840 meshes, 1449-1505 skins):
//  - one Mesh per glTF mesh, one MeshPart per primitive, in order; each
//    primitive's vertices appended after the ones before it (Mesh.AddVertices,
//    mesh.go:678-711) and its indices offset by the part's first vertex
//    (MeshPart.AddTriangles, mesh.go:2274-2314);
//  - POSITION and NORMAL copied, TEXCOORD_0 as (u, -(v - 1)) in float32;
//  - COLOR_0 read as 8 bits per channel, divided by 255 in float32 and
//    converted to sRGB (Color4.ConvertTosRGB, color.go:194-214) with
//    Pow(x, 1/2.4) in float64 (math32.Pow); a vertex without COLOR_0 is white
//    (mesh.go:543-556); the colours are the mesh's active channel only when the
//    mesh extras name its colour channels (t3dVertexColorNames__, gltf.go:799-
//    837), else the mesh has no active channel;
//  - JOINTS_0 / WEIGHTS_0: an influence is kept only when its weight is > 0
//    (gltf.go:671-680); here the four slots keep their places and a dropped
//    influence gets weight 0, which every reader of the pool skips;
//  - morph targets as shape keys (gltf.go:747-790) for single-primitive meshes;
//  - Dimensions: the min and max over the positions, from ±MaxFloat32
//    (Mesh.UpdateDimensions, mesh.go:719-750);
//  - skins: the joint node indices, and each inverse bind matrix with
//    column r of the Matrix4 set from the r-th four floats (gltf.go:1487-1497).
// Buffers come from the GLB binary chunk or base64 data URIs; sparse
// accessors, external files and meshes the reference would panic on (no
// indices, index counts not a multiple of three, 0 vertices) are refused.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/t3d.h"

namespace {

// ---- a small JSON reader (RFC 8259 values into a tree) --------------------

struct Json {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  bool b = false;
  double n = 0;
  std::string s;
  std::vector<Json> a;
  std::vector<std::pair<std::string, Json>> o;
  const Json* get(const char* key) const {
    if (kind != Obj) return nullptr;
    for (auto& kv : o)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

struct JsonParser {
  const char* p;
  const char* end;
  std::string err;
  int depth = 0;
  void ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
  }
  bool fail(const char* m) {
    if (err.empty()) err = m;
    return false;
  }
  bool lit(const char* w) {
    size_t n = std::strlen(w);
    if ((size_t)(end - p) < n || std::memcmp(p, w, n) != 0) return fail("json: bad literal");
    p += n;
    return true;
  }
  static void utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) out += (char)cp;
    else if (cp < 0x800) {
      out += (char)(0xC0 | (cp >> 6));
      out += (char)(0x80 | (cp & 0x3F));
    } else if (cp < 0x10000) {
      out += (char)(0xE0 | (cp >> 12));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    } else {
      out += (char)(0xF0 | (cp >> 18));
      out += (char)(0x80 | ((cp >> 12) & 0x3F));
      out += (char)(0x80 | ((cp >> 6) & 0x3F));
      out += (char)(0x80 | (cp & 0x3F));
    }
  }
  bool hex4(uint32_t& v) {
    if (end - p < 4) return fail("json: short \\u escape");
    v = 0;
    for (int i = 0; i < 4; i++) {
      char c = *p++;
      v <<= 4;
      if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
      else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
      else return fail("json: bad \\u escape");
    }
    return true;
  }
  bool str(std::string& out) {
    if (p >= end || *p != '"') return fail("json: expected a string");
    p++;
    while (p < end && *p != '"') {
      if ((unsigned char)*p < 0x20) return fail("json: control character in a string");
      if (*p != '\\') {
        out += *p++;
        continue;
      }
      if (++p >= end) return fail("json: bad escape");
      char c = *p++;
      switch (c) {
        case '"': out += '"'; break;
        case '\\': out += '\\'; break;
        case '/': out += '/'; break;
        case 'b': out += '\b'; break;
        case 'f': out += '\f'; break;
        case 'n': out += '\n'; break;
        case 'r': out += '\r'; break;
        case 't': out += '\t'; break;
        case 'u': {
          uint32_t cp;
          if (!hex4(cp)) return false;
          if (cp >= 0xD800 && cp < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
            p += 2;
            uint32_t lo;
            if (!hex4(lo)) return false;
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(out, cp);
          break;
        }
        default: return fail("json: bad escape");
      }
    }
    if (p >= end) return fail("json: unterminated string");
    p++;
    return true;
  }
  bool value(Json& v) {
    if (++depth > 256) return fail("json: nesting too deep");
    ws();
    if (p >= end) return fail("json: unexpected end");
    bool ok = true;
    if (*p == '{') {
      v.kind = Json::Obj;
      p++;
      ws();
      if (p < end && *p == '}') p++;
      else
        for (;;) {
          std::pair<std::string, Json> kv;
          ws();
          if (!str(kv.first)) return false;
          ws();
          if (p >= end || *p != ':') return fail("json: expected ':'");
          p++;
          if (!value(kv.second)) return false;
          v.o.push_back(std::move(kv));
          ws();
          if (p < end && *p == ',') { p++; continue; }
          if (p < end && *p == '}') { p++; break; }
          return fail("json: expected ',' or '}'");
        }
    } else if (*p == '[') {
      v.kind = Json::Arr;
      p++;
      ws();
      if (p < end && *p == ']') p++;
      else
        for (;;) {
          v.a.emplace_back();
          if (!value(v.a.back())) return false;
          ws();
          if (p < end && *p == ',') { p++; continue; }
          if (p < end && *p == ']') { p++; break; }
          return fail("json: expected ',' or ']'");
        }
    } else if (*p == '"') {
      v.kind = Json::Str;
      ok = str(v.s);
    } else if (*p == 't') {
      v.kind = Json::Bool;
      v.b = true;
      ok = lit("true");
    } else if (*p == 'f') {
      v.kind = Json::Bool;
      ok = lit("false");
    } else if (*p == 'n') {
      ok = lit("null");
    } else {
      const char* s0 = p;
      if (p < end && *p == '-') p++;
      while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-')) p++;
      if (p == s0) return fail("json: unexpected character");
      std::string num(s0, p);
      char* e = nullptr;
      v.kind = Json::Num;
      v.n = std::strtod(num.c_str(), &e);
      if (!e || *e) return fail("json: bad number");
    }
    depth--;
    return ok;
  }
};

// ---- glTF document ---------------------------------------------------------

bool as_uint(const Json* j, uint64_t& out) {
  if (!j || j->kind != Json::Num || j->n < 0 || j->n != std::floor(j->n) || j->n > 9.0e15) return false;
  out = (uint64_t)j->n;
  return true;
}
uint64_t uint_or(const Json* j, uint64_t dflt) {
  uint64_t v;
  return as_uint(j, v) ? v : dflt;
}

int b64val(char c) {
  if (c >= 'A' && c <= 'Z') return c - 'A';
  if (c >= 'a' && c <= 'z') return c - 'a' + 26;
  if (c >= '0' && c <= '9') return c - '0' + 52;
  if (c == '+') return 62;
  if (c == '/') return 63;
  return -1;
}

struct Accessor {
  uint64_t count = 0;
  uint32_t comp = 0;       // componentType
  uint32_t ncomp = 0;      // components per element
  bool normalized = false;
  const uint8_t* base = nullptr;   // first element
  uint64_t stride = 0;
  bool zero = false;       // no bufferView: all zeros
};

uint32_t comp_size(uint32_t ct) {
  switch (ct) {
    case 5120: case 5121: return 1;
    case 5122: case 5123: return 2;
    case 5125: case 5126: return 4;
    default: return 0;
  }
}

// One component of an accessor element as the glTF spec defines its value:
// float as is; normalized integers scaled to [0, 1] (unsigned) or [-1, 1]
// (signed); other integers as their value.
float comp_float(const Accessor& a, uint64_t i, uint32_t c) {
  if (a.zero) return 0.0f;
  const uint8_t* q = a.base + i * a.stride + (uint64_t)c * comp_size(a.comp);
  switch (a.comp) {
    case 5126: { float f; std::memcpy(&f, q, 4); return f; }
    case 5121: { uint8_t v = *q; return a.normalized ? (float)v / 255.0f : (float)v; }
    case 5120: { int8_t v; std::memcpy(&v, q, 1); return a.normalized ? std::fmax((float)v / 127.0f, -1.0f) : (float)v; }
    case 5123: { uint16_t v; std::memcpy(&v, q, 2); return a.normalized ? (float)v / 65535.0f : (float)v; }
    case 5122: { int16_t v; std::memcpy(&v, q, 2); return a.normalized ? std::fmax((float)v / 32767.0f, -1.0f) : (float)v; }
    case 5125: { uint32_t v; std::memcpy(&v, q, 4); return (float)v; }
  }
  return 0.0f;
}
uint32_t comp_uint(const Accessor& a, uint64_t i, uint32_t c) {
  if (a.zero) return 0;
  const uint8_t* q = a.base + i * a.stride + (uint64_t)c * comp_size(a.comp);
  switch (a.comp) {
    case 5121: return *q;
    case 5123: { uint16_t v; std::memcpy(&v, q, 2); return v; }
    case 5125: { uint32_t v; std::memcpy(&v, q, 4); return v; }
  }
  return 0;
}
// A colour component as the 8 bits the reference keeps: 8-bit as is; 16-bit
// and float by the glTF quantisation round(x * 255) (the reference reads them
// through its glTF library's colour reader, which is not part of it).
uint8_t comp_u8(const Accessor& a, uint64_t i, uint32_t c) {
  if (a.zero) return 0;
  if (a.comp == 5121) return (uint8_t)comp_uint(a, i, c);
  if (a.comp == 5123) return (uint8_t)std::lround((double)comp_uint(a, i, c) / 257.0);
  float f = comp_float(a, i, c);
  if (!(f > 0.0f)) return 0;
  if (f >= 1.0f) return 255;
  return (uint8_t)std::lround((double)f * 255.0);
}

// Color4.ConvertTosRGB (color.go:194-214) of one channel x = c / 255 in float32.
float srgb_channel(float x) {
  if (x <= 0.0031308f) return x * 12.92f;
  const float p = (float)std::pow((double)x, (double)(float)(5.0 / 12.0));   // math32.Pow(x, 1/2.4)
  const float a = 1.055f * p;
  return a - 0.055f;
}

struct MeshOut {
  std::string name;
  std::vector<float> pos, nrm, col, uv, weights, skp, skn;
  std::vector<uint16_t> bones;
  std::vector<uint32_t> idx;
  std::vector<t3d_part_range> parts;
  std::vector<int32_t> part_material;
  uint32_t K = 0;
  bool has_colors = false, skinned = false;
  float dmin[3], dmax[3];
  t3d_mesh_desc desc{};
};

struct SkinOut {
  std::string name;
  std::vector<int32_t> joints;
  std::vector<t3d_mat4> inverse_bind;
};

}  // namespace

struct t3d_gltf {
  std::vector<uint8_t> bin;               // the GLB binary chunk
  std::vector<std::vector<uint8_t>> uris; // decoded data URIs, per buffer
  std::vector<const uint8_t*> buf_ptr;
  std::vector<uint64_t> buf_len;
  Json doc;
  std::vector<MeshOut> meshes;
  std::vector<SkinOut> skins;
  std::string err;
};

namespace {

bool gerr(t3d_gltf* g, const std::string& m) {
  if (g->err.empty()) g->err = m;
  return false;
}

const Json* arr_at(const Json* arr, uint64_t i) {
  if (!arr || arr->kind != Json::Arr || i >= arr->a.size()) return nullptr;
  return &arr->a[i];
}

bool load_buffers(t3d_gltf* g) {
  const Json* bufs = g->doc.get("buffers");
  const size_t n = bufs && bufs->kind == Json::Arr ? bufs->a.size() : 0;
  g->uris.resize(n);
  g->buf_ptr.assign(n, nullptr);
  g->buf_len.assign(n, 0);
  for (size_t i = 0; i < n; i++) {
    const Json& b = bufs->a[i];
    uint64_t len;
    if (!as_uint(b.get("byteLength"), len)) return gerr(g, "gltf: buffer " + std::to_string(i) + " has no byteLength");
    const Json* uri = b.get("uri");
    if (!uri) {   // the GLB binary chunk
      if (i != 0 || g->bin.size() < len) return gerr(g, "gltf: buffer " + std::to_string(i) + " has no data");
      g->buf_ptr[i] = g->bin.data();
      g->buf_len[i] = len;
      continue;
    }
    if (uri->kind != Json::Str) return gerr(g, "gltf: buffer uri is not a string");
    const std::string& u = uri->s;
    const size_t comma = u.find(',');
    if (u.compare(0, 5, "data:") != 0 || comma == std::string::npos || u.find(";base64") == std::string::npos ||
        u.find(";base64") > comma)
      return gerr(g, "gltf: buffer " + std::to_string(i) + " is an external file (only GLB and base64 data URIs are read)");
    std::vector<uint8_t>& out = g->uris[i];
    uint32_t acc = 0;
    int nb = 0;
    for (size_t k = comma + 1; k < u.size(); k++) {
      if (u[k] == '=') break;
      const int v = b64val(u[k]);
      if (v < 0) return gerr(g, "gltf: bad base64 in buffer " + std::to_string(i));
      acc = (acc << 6) | (uint32_t)v;
      nb += 6;
      if (nb >= 8) {
        nb -= 8;
        out.push_back((uint8_t)(acc >> nb));
        acc &= (1u << nb) - 1;
      }
    }
    if (out.size() < len) return gerr(g, "gltf: buffer " + std::to_string(i) + " is shorter than its byteLength");
    g->buf_ptr[i] = out.data();
    g->buf_len[i] = len;
  }
  return true;
}

bool read_accessor(t3d_gltf* g, uint64_t index, Accessor& a, const char* what) {
  const Json* acc = arr_at(g->doc.get("accessors"), index);
  if (!acc) return gerr(g, std::string("gltf: ") + what + " names a missing accessor");
  if (acc->get("sparse")) return gerr(g, std::string("gltf: ") + what + " uses a sparse accessor");
  uint64_t ct, count;
  if (!as_uint(acc->get("componentType"), ct) || !comp_size((uint32_t)ct)) return gerr(g, std::string("gltf: ") + what + ": bad componentType");
  if (!as_uint(acc->get("count"), count)) return gerr(g, std::string("gltf: ") + what + ": bad count");
  const Json* type = acc->get("type");
  if (!type || type->kind != Json::Str) return gerr(g, std::string("gltf: ") + what + ": no type");
  static const std::pair<const char*, uint32_t> types[] = {{"SCALAR", 1}, {"VEC2", 2}, {"VEC3", 3}, {"VEC4", 4},
                                                            {"MAT2", 4},   {"MAT3", 9}, {"MAT4", 16}};
  a = Accessor{};
  for (auto& t : types)
    if (type->s == t.first) a.ncomp = t.second;
  if (!a.ncomp) return gerr(g, std::string("gltf: ") + what + ": bad type");
  a.comp = (uint32_t)ct;
  a.count = count;
  const Json* norm = acc->get("normalized");
  a.normalized = norm && norm->kind == Json::Bool && norm->b;
  const uint64_t esize = (uint64_t)a.ncomp * comp_size(a.comp);
  uint64_t bv;
  if (!as_uint(acc->get("bufferView"), bv)) {
    a.zero = true;
    return true;
  }
  const Json* view = arr_at(g->doc.get("bufferViews"), bv);
  if (!view) return gerr(g, std::string("gltf: ") + what + " names a missing bufferView");
  uint64_t buf, vlen;
  if (!as_uint(view->get("buffer"), buf) || buf >= g->buf_ptr.size() || !as_uint(view->get("byteLength"), vlen))
    return gerr(g, std::string("gltf: ") + what + ": bad bufferView");
  const uint64_t voff = uint_or(view->get("byteOffset"), 0);
  const uint64_t aoff = uint_or(acc->get("byteOffset"), 0);
  a.stride = uint_or(view->get("byteStride"), 0);
  if (!a.stride) a.stride = esize;
  if (a.stride < esize) return gerr(g, std::string("gltf: ") + what + ": byteStride smaller than an element");
  if (voff > g->buf_len[buf] || vlen > g->buf_len[buf] - voff) return gerr(g, std::string("gltf: ") + what + ": bufferView outside its buffer");
  if (count) {
    const uint64_t need = aoff + (count - 1) * a.stride + esize;
    if (aoff > vlen || need > vlen || (count - 1) > UINT64_MAX / a.stride)
      return gerr(g, std::string("gltf: ") + what + ": accessor outside its bufferView");
  }
  a.base = g->buf_ptr[buf] + voff + aoff;
  return true;
}

bool attr(const Json* attrs, const char* name, uint64_t& index) {
  return attrs && as_uint(attrs->get(name), index);
}

bool build_mesh(t3d_gltf* g, const Json& m, MeshOut& out) {
  if (const Json* nm = m.get("name"))
    if (nm->kind == Json::Str) out.name = nm->s;
  const Json* prims = m.get("primitives");
  if (!prims || prims->kind != Json::Arr || prims->a.empty()) return gerr(g, "gltf: mesh " + out.name + " has no primitives");
  const Json* extras = m.get("extras");
  const Json* vcn = extras ? extras->get("t3dVertexColorNames__") : nullptr;
  if (const Json* as = extras ? extras->get("t3dAutoSubdivide__") : nullptr)
    if (as->kind == Json::Num && as->n > 0)   // subdivided again at render time (gltf.go:560-562, 795-797)
      return gerr(g, "gltf: mesh " + out.name + " is auto-subdivided");
  out.has_colors = vcn && vcn->kind == Json::Arr && !vcn->a.empty();
  bool any_targets = false;
  for (const Json& p : prims->a)
    if (const Json* t = p.get("targets"))
      if (t->kind == Json::Arr && !t->a.empty()) any_targets = true;
  if (any_targets && prims->a.size() > 1)
    return gerr(g, "gltf: mesh " + out.name + " has morph targets on more than one primitive");
  uint64_t V = 0, T = 0;
  for (size_t pi = 0; pi < prims->a.size(); pi++) {
    const Json& p = prims->a[pi];
    const Json* attrs = p.get("attributes");
    uint64_t ia;
    Accessor pa, na, ua, ca, ja, wa, xa;
    if (!attr(attrs, "POSITION", ia) || !read_accessor(g, ia, pa, "POSITION")) return gerr(g, "gltf: a primitive has no POSITION");
    if (pa.ncomp != 3 || pa.comp != 5126) return gerr(g, "gltf: POSITION is not float VEC3");
    const uint64_t n = pa.count;
    if (n == 0) return gerr(g, "gltf: a primitive has no vertices");   // Mesh.AddVertices panics (mesh.go:680-682)
    const bool hn = attr(attrs, "NORMAL", ia);
    if (hn && (!read_accessor(g, ia, na, "NORMAL") || na.ncomp != 3 || na.count < n)) return gerr(g, "gltf: bad NORMAL");
    const bool hu = attr(attrs, "TEXCOORD_0", ia);
    if (hu && (!read_accessor(g, ia, ua, "TEXCOORD_0") || ua.ncomp != 2 || ua.count < n)) return gerr(g, "gltf: bad TEXCOORD_0");
    const bool hc = attr(attrs, "COLOR_0", ia);
    if (hc && (!read_accessor(g, ia, ca, "COLOR_0") || (ca.ncomp != 3 && ca.ncomp != 4) || ca.count < n ||
               (ca.comp != 5121 && ca.comp != 5123 && ca.comp != 5126)))
      return gerr(g, "gltf: bad COLOR_0");
    const bool hw = attr(attrs, "WEIGHTS_0", ia);
    if (hw) {
      uint64_t ij;
      if (!read_accessor(g, ia, wa, "WEIGHTS_0") || wa.ncomp != 4 || wa.count < n) return gerr(g, "gltf: bad WEIGHTS_0");
      if (!attr(attrs, "JOINTS_0", ij) || !read_accessor(g, ij, ja, "JOINTS_0") || ja.ncomp != 4 || ja.count < n ||
          (ja.comp != 5121 && ja.comp != 5123))
        return gerr(g, "gltf: WEIGHTS_0 without valid JOINTS_0");
      out.skinned = true;
    }
    uint64_t ii;
    if (!as_uint(p.get("indices"), ii) || !read_accessor(g, ii, xa, "indices") || xa.ncomp != 1 ||
        (xa.comp != 5121 && xa.comp != 5123 && xa.comp != 5125))
      return gerr(g, "gltf: a primitive has no usable indices");
    if (xa.count == 0 || xa.count % 3) return gerr(g, "gltf: a primitive's index count is not a multiple of 3");
    // vertices (Mesh.AddVertices)
    for (uint64_t i = 0; i < n; i++) {
      for (uint32_t c = 0; c < 3; c++) {
        out.pos.push_back(comp_float(pa, i, c));
        out.nrm.push_back(hn ? comp_float(na, i, c) : 0.0f);
      }
      const float u = hu ? comp_float(ua, i, 0) : 0.0f;
      const float v = hu ? comp_float(ua, i, 1) : 0.0f;
      out.uv.push_back(u);
      out.uv.push_back(hu ? -(v - 1.0f) : 0.0f);   // gltf.go:610-611
      if (hc) {
        for (uint32_t c = 0; c < 3; c++) out.col.push_back(srgb_channel((float)comp_u8(ca, i, c) / 255.0f));
        out.col.push_back(ca.ncomp == 4 ? (float)comp_u8(ca, i, 3) / 255.0f : 1.0f);
      } else {
        for (int c = 0; c < 4; c++) out.col.push_back(1.0f);
      }
      for (uint32_t k = 0; k < 4; k++) {
        const float w = hw ? comp_float(wa, i, k) : 0.0f;
        const bool keep = w > 0.0f;
        out.weights.push_back(keep ? w : 0.0f);
        out.bones.push_back(keep ? (uint16_t)comp_uint(ja, i, k) : (uint16_t)0);
      }
    }
    // indices (MeshPart.AddTriangles)
    t3d_part_range r{(uint32_t)T, (uint32_t)(xa.count / 3)};
    for (uint64_t k = 0; k < xa.count; k++) {
      const uint64_t id = comp_uint(xa, k, 0);
      if (id >= n) return gerr(g, "gltf: an index is past its primitive's vertices");
      out.idx.push_back((uint32_t)(id + V));
    }
    out.parts.push_back(r);
    uint64_t mat;
    out.part_material.push_back(as_uint(p.get("material"), mat) ? (int32_t)mat : -1);
    // morph targets (gltf.go:747-790)
    if (const Json* tg = p.get("targets"))
      if (tg->kind == Json::Arr)
        for (const Json& t : tg->a) {
          std::vector<float> kp(3 * n, 0.0f), kn(3 * n, 0.0f);
          Accessor ta;
          if (attr(&t, "POSITION", ia)) {
            if (!read_accessor(g, ia, ta, "target POSITION") || ta.ncomp != 3 || ta.count < n) return gerr(g, "gltf: bad target POSITION");
            for (uint64_t i = 0; i < n; i++)
              for (uint32_t c = 0; c < 3; c++) kp[3 * i + c] = comp_float(ta, i, c);
          }
          if (attr(&t, "NORMAL", ia)) {
            if (!read_accessor(g, ia, ta, "target NORMAL") || ta.ncomp != 3 || ta.count < n) return gerr(g, "gltf: bad target NORMAL");
            for (uint64_t i = 0; i < n; i++)
              for (uint32_t c = 0; c < 3; c++) kn[3 * i + c] = comp_float(ta, i, c);
          }
          out.skp.insert(out.skp.end(), kp.begin(), kp.end());
          out.skn.insert(out.skn.end(), kn.begin(), kn.end());
          out.K++;
        }
    V += n;
    T += r.triangle_count;
    if (V > 0xFFFFFFFFull || T > 0x3FFFFFFFull) return gerr(g, "gltf: mesh " + out.name + " is too large");
  }
  // Mesh.UpdateDimensions (mesh.go:719-750)
  const float big = 3.40282346638528859811704183484516925440e+38f;
  for (int c = 0; c < 3; c++) {
    out.dmin[c] = big;
    out.dmax[c] = -big;
  }
  for (uint64_t i = 0; i < V; i++)
    for (int c = 0; c < 3; c++) {
      const float x = out.pos[3 * i + c];
      if (out.dmin[c] > x) out.dmin[c] = x;
      if (out.dmax[c] < x) out.dmax[c] = x;
    }
  t3d_mesh_desc& d = out.desc;
  d = t3d_mesh_desc{};
  d.vertex_count = (uint32_t)V;
  d.positions = out.pos.data();
  d.normals = out.nrm.data();
  d.colors = out.has_colors ? out.col.data() : nullptr;
  d.uvs = out.uv.data();
  d.bone_indices = out.skinned ? out.bones.data() : nullptr;
  d.bone_weights = out.skinned ? out.weights.data() : nullptr;
  d.shape_key_count = out.K;
  d.shape_key_positions = out.K ? out.skp.data() : nullptr;
  d.shape_key_normals = out.K ? out.skn.data() : nullptr;
  d.triangle_count = (uint32_t)T;
  d.indices = out.idx.data();
  d.part_count = (uint32_t)out.parts.size();
  d.parts = out.parts.data();
  for (int c = 0; c < 3; c++) {
    d.dimensions_min[c] = out.dmin[c];
    d.dimensions_max[c] = out.dmax[c];
  }
  return true;
}

bool build_skin(t3d_gltf* g, const Json& s, SkinOut& out) {
  if (const Json* nm = s.get("name"))
    if (nm->kind == Json::Str) out.name = nm->s;
  const Json* joints = s.get("joints");
  if (!joints || joints->kind != Json::Arr) return gerr(g, "gltf: skin " + out.name + " has no joints");
  const size_t nodes = g->doc.get("nodes") && g->doc.get("nodes")->kind == Json::Arr ? g->doc.get("nodes")->a.size() : 0;
  for (const Json& j : joints->a) {
    uint64_t v;
    if (!as_uint(&j, v) || v >= nodes) return gerr(g, "gltf: skin " + out.name + " names a missing joint node");
    out.joints.push_back((int32_t)v);
  }
  out.inverse_bind.assign(out.joints.size(), t3d_mat4{});
  for (size_t b = 0; b < out.joints.size(); b++)
    for (int k = 0; k < 4; k++) out.inverse_bind[b].m[k][k] = 1.0f;
  uint64_t ia;
  if (as_uint(s.get("inverseBindMatrices"), ia)) {
    Accessor a;
    if (!read_accessor(g, ia, a, "inverseBindMatrices") || a.ncomp != 16 || a.comp != 5126 || a.count < out.joints.size())
      return gerr(g, "gltf: skin " + out.name + ": bad inverseBindMatrices");
    for (size_t b = 0; b < out.joints.size(); b++)
      for (int col = 0; col < 4; col++)        // Matrix4.SetColumn(col, the col-th four floats)
        for (int row = 0; row < 4; row++) out.inverse_bind[b].m[row][col] = comp_float(a, b, (uint32_t)(4 * col + row));
  }
  return true;
}

bool parse(t3d_gltf* g, const uint8_t* data, uint64_t size) {
  const char* json = reinterpret_cast<const char*>(data);
  uint64_t json_len = size;
  if (size >= 12 && std::memcmp(data, "glTF", 4) == 0) {   // GLB container
    uint32_t version, length;
    std::memcpy(&version, data + 4, 4);
    std::memcpy(&length, data + 8, 4);
    if (version != 2 || length > size || length < 20) return gerr(g, "gltf: bad GLB header");
    uint64_t off = 12;
    json = nullptr;
    while (off + 8 <= length) {
      uint32_t clen, ctype;
      std::memcpy(&clen, data + off, 4);
      std::memcpy(&ctype, data + off + 4, 4);
      off += 8;
      if (clen > length - off) return gerr(g, "gltf: GLB chunk past the end");
      if (ctype == 0x4E4F534Au && !json) {
        json = reinterpret_cast<const char*>(data + off);
        json_len = clen;
      } else if (ctype == 0x004E4942u && g->bin.empty()) {
        g->bin.assign(data + off, data + off + clen);
      }
      off += (clen + 3) & ~3ull;
    }
    if (!json) return gerr(g, "gltf: GLB without a JSON chunk");
  }
  JsonParser jp{json, json + json_len};
  if (!jp.value(g->doc)) return gerr(g, jp.err);
  if (g->doc.kind != Json::Obj) return gerr(g, "gltf: the document is not an object");
  if (!load_buffers(g)) return false;
  if (const Json* ms = g->doc.get("meshes"))
    if (ms->kind == Json::Arr) {
      g->meshes.resize(ms->a.size());
      for (size_t i = 0; i < ms->a.size(); i++)
        if (!build_mesh(g, ms->a[i], g->meshes[i])) return false;
    }
  if (const Json* ss = g->doc.get("skins"))
    if (ss->kind == Json::Arr) {
      g->skins.resize(ss->a.size());
      for (size_t i = 0; i < ss->a.size(); i++)
        if (!build_skin(g, ss->a[i], g->skins[i])) return false;
    }
  return true;
}

void copy_err(const std::string& e, char* out, uint32_t len) {
  if (!out || !len) return;
  const size_t n = std::min<size_t>(e.size(), len - 1);
  std::memcpy(out, e.data(), n);
  out[n] = 0;
}

}  // namespace

extern "C" {

t3d_status t3d_gltf_parse(const void* data, uint64_t size, t3d_gltf** out, char* err, uint32_t err_len) {
  if (!out) return T3D_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (!data || !size) {
    copy_err("gltf: no data", err, err_len);
    return T3D_ERR_INVALID_ARGUMENT;
  }
  auto g = std::make_unique<t3d_gltf>();
  if (!parse(g.get(), static_cast<const uint8_t*>(data), size)) {
    copy_err(g->err, err, err_len);
    return T3D_ERR_INVALID_ARGUMENT;
  }
  copy_err("", err, err_len);
  *out = g.release();
  return T3D_OK;
}

void t3d_gltf_free(t3d_gltf* g) { delete g; }

uint32_t t3d_gltf_mesh_count(const t3d_gltf* g) { return g ? (uint32_t)g->meshes.size() : 0; }

t3d_status t3d_gltf_mesh(const t3d_gltf* g, uint32_t i, t3d_gltf_mesh_info* info, t3d_mesh_desc* desc,
                         const int32_t** part_materials) {
  if (!g || i >= g->meshes.size()) return T3D_ERR_NOT_FOUND;
  const MeshOut& m = g->meshes[i];
  if (info) {
    *info = t3d_gltf_mesh_info{};
    const size_t n = std::min<size_t>(m.name.size(), sizeof info->name - 1);
    std::memcpy(info->name, m.name.data(), n);
    info->vertex_count = m.desc.vertex_count;
    info->triangle_count = m.desc.triangle_count;
    info->part_count = m.desc.part_count;
    info->shape_key_count = m.K;
    info->has_colors = m.has_colors ? 1 : 0;
    info->skinned = m.skinned ? 1 : 0;
  }
  if (desc) *desc = m.desc;
  if (part_materials) *part_materials = m.part_material.data();
  return T3D_OK;
}

uint32_t t3d_gltf_skin_count(const t3d_gltf* g) { return g ? (uint32_t)g->skins.size() : 0; }

t3d_status t3d_gltf_skin(const t3d_gltf* g, uint32_t i, uint32_t* joint_count, const int32_t** joints,
                         const t3d_mat4** inverse_bind) {
  if (!g || i >= g->skins.size()) return T3D_ERR_NOT_FOUND;
  const SkinOut& s = g->skins[i];
  if (joint_count) *joint_count = (uint32_t)s.joints.size();
  if (joints) *joints = s.joints.data();
  if (inverse_bind) *inverse_bind = s.inverse_bind.data();
  return T3D_OK;
}

t3d_status t3d_gltf_upload(t3d_ctx* ctx, const t3d_gltf* g, uint64_t* handles) {
  if (!ctx || !g || (!handles && !g->meshes.empty())) return T3D_ERR_INVALID_ARGUMENT;
  for (size_t i = 0; i < g->meshes.size(); i++) handles[i] = 0;
  for (size_t i = 0; i < g->meshes.size(); i++) {
    t3d_status st = t3d_mesh_upload(ctx, &g->meshes[i].desc, &handles[i]);
    if (st) {   // nothing of a partial load stays resident
      for (size_t k = 0; k < i; k++) (void)t3d_mesh_free(ctx, handles[k]);
      for (size_t k = 0; k < g->meshes.size(); k++) handles[k] = 0;
      return st;
    }
  }
  return T3D_OK;
}

}  // extern "C"
