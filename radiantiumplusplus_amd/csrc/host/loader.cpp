// Scene JSON loading: BuildContext::SetFromJson + Build
// (module/rad.offline/src/build/build_context.cpp:47-321) producing a
// rad_scene_desc, with the asset readers it relies on: Wavefront OBJ
// (module/rad.core/src/wavefront_obj_reader.cpp:37-463) and mitsuba VOL v3
// (module/rad.core/src/volume_reader.cpp:65-124).
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <set>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/rad_hip.h"
#include "image.h"
#include "json.h"

namespace rad {
void SetLastError(const std::string& msg);
}

namespace {

struct LoadError : std::runtime_error {
  rad_status code;
  LoadError(rad_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

using rad::json::Value;

// ------------------------------------------------------------------ matrices
struct M4 {
  float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  M4 operator*(const M4& o) const {
    M4 r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        float s = 0;
        for (int k = 0; k < 4; k++) s += m[i * 4 + k] * o.m[k * 4 + j];
        r.m[i * 4 + j] = s;
      }
    return r;
  }
};

float Num(const Value& v, const char* what) {
  if (!v.is_number()) throw LoadError(RAD_ERR_INVALID_ARG, std::string("node type is not number: ") + what);
  return float(v.num);
}

// JsonToVector (core/include/rad/core/config_node.h:12-34): array of N numbers or one number.
template <int N>
void Vec(const Value& v, float* out, const char* what) {
  if (v.is_array()) {
    if (v.arr.size() != N) throw LoadError(RAD_ERR_INVALID_ARG, std::string("node array size must be ") + std::to_string(N) + ": " + what);
    for (int i = 0; i < N; i++) out[i] = Num(v.arr[i], what);
  } else if (v.is_number()) {
    for (int i = 0; i < N; i++) out[i] = float(v.num);
  } else {
    throw LoadError(RAD_ERR_INVALID_ARG, std::string("node must be array or number: ") + what);
  }
}

template <int N>
void VecOr(const Value& node, const char* key, float* out, std::initializer_list<float> def) {
  const Value* v = node.find(key);
  if (!v) {
    int i = 0;
    for (float d : def) out[i++] = d;
    return;
  }
  Vec<N>(*v, out, key);
}

float FloatOr(const Value& node, const char* key, float def) {
  const Value* v = node.find(key);
  return v ? Num(*v, key) : def;
}

int IntOr(const Value& node, const char* key, int def) {
  const Value* v = node.find(key);
  return v ? int(Num(*v, key)) : def;
}

bool BoolOr(const Value& node, const char* key, bool def) {
  const Value* v = node.find(key);
  if (!v) return def;
  if (!v->is_bool()) throw LoadError(RAD_ERR_INVALID_ARG, std::string("node type is not bool: ") + key);
  return v->b;
}

std::string StrOr(const Value& node, const char* key, const std::string& def) {
  const Value* v = node.find(key);
  if (!v) return def;
  if (!v->is_string()) throw LoadError(RAD_ERR_INVALID_ARG, std::string("node type is not string: ") + key);
  return v->str;
}

std::string Str(const Value& node, const char* key) {
  const Value* v = node.find(key);
  if (!v) throw LoadError(RAD_ERR_INVALID_ARG, std::string("no child: ") + key);
  if (!v->is_string()) throw LoadError(RAD_ERR_INVALID_ARG, std::string("node type is not string: ") + key);
  return v->str;
}

// Eigen::AngleAxis(Radian(angle), axis).toRotationMatrix()
void AxisAngle(const float* axis, float angleDeg, float R[9]) {
  float a = angleDeg * (3.14159265358979323846f / 180.f);
  float s = std::sin(a), c = std::cos(a), t = 1.f - c;
  float x = axis[0], y = axis[1], z = axis[2];
  float r[9] = {t * x * x + c,     t * x * y - s * z, t * x * z + s * y, t * x * y + s * z, t * y * y + c,
                t * y * z - s * x, t * x * z - s * y, t * y * z + s * x, t * z * z + c};
  std::memcpy(R, r, sizeof(r));
}

// ConfigNodeAsRotate (config_node_ext.cpp:7-28)
void AsRotate(const Value& n, float R[9]) {
  if (n.is_object()) {
    float axis[3], angle;
    VecOr<3>(n, "axis", axis, {0, 1, 0});
    angle = FloatOr(n, "angle", 0.f);
    AxisAngle(axis, angle, R);
  } else if (n.is_array() && n.arr.size() == 9) {
    for (int i = 0; i < 9; i++) R[i] = Num(n.arr[i], "rotate");
  } else if (n.is_array() && n.arr.size() == 4) {
    // Eigen::Quaternion(Vector4) takes (x, y, z, w) coefficient order
    float x = Num(n.arr[0], "q"), y = Num(n.arr[1], "q"), z = Num(n.arr[2], "q"), w = Num(n.arr[3], "q");
    float r[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - z * w),     2 * (x * z + y * w),
                  2 * (x * y + z * w),     1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                  2 * (x * z - y * w),     2 * (y * z + x * w),     1 - 2 * (x * x + y * y)};
    std::memcpy(R, r, sizeof(r));
  } else {
    throw LoadError(RAD_ERR_INVALID_ARG, "rotate must be {axis, angle}, a 3x3 matrix or a quaternion");
  }
}

// ConfigNodeAsTransform (config_node_ext.cpp:52-71)
M4 AsTransform(const Value& n) {
  M4 r;
  if (n.is_array()) {
    if (n.arr.size() != 16) throw LoadError(RAD_ERR_INVALID_ARG, "to_world array must hold 16 numbers");
    for (int i = 0; i < 16; i++) r.m[i] = Num(n.arr[i], "to_world");
    return r;
  }
  if (!n.is_object()) throw LoadError(RAD_ERR_INVALID_ARG, "transform node must be mat4 or object");
  float t[3], s[3], R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  VecOr<3>(n, "translate", t, {0, 0, 0});
  VecOr<3>(n, "scale", s, {1, 1, 1});
  if (const Value* rot = n.find("rotate")) AsRotate(*rot, R);
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) r.m[i * 4 + j] = R[i * 3 + j] * s[j];
    r.m[i * 4 + 3] = t[i];
  }
  r.m[12] = r.m[13] = r.m[14] = 0;
  r.m[15] = 1;
  return r;
}

// math_ext.h:65-75
M4 LookAtLeftHand(const float* o, const float* t, const float* up) {
  auto norm = [](float* v) {
    float n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n > 0)
      for (int i = 0; i < 3; i++) v[i] /= n;
  };
  auto cross = [](const float* a, const float* b, float* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
  };
  float dir[3] = {t[0] - o[0], t[1] - o[1], t[2] - o[2]};
  norm(dir);
  float u[3] = {up[0], up[1], up[2]};
  norm(u);
  float left[3], realUp[3];
  cross(u, dir, left);
  norm(left);
  cross(dir, left, realUp);
  norm(realUp);
  M4 r;
  for (int i = 0; i < 3; i++) {
    r.m[i * 4 + 0] = left[i];
    r.m[i * 4 + 1] = realUp[i];
    r.m[i * 4 + 2] = dir[i];
    r.m[i * 4 + 3] = o[i];
  }
  return r;
}

// ------------------------------------------------------------------ assets
struct Model {
  std::vector<float> pos, nrm, uv;
  std::vector<uint32_t> idx;
};

struct ObjFace {
  int v[3] = {0, 0, 0}, vt[3] = {0, 0, 0}, vn[3] = {0, 0, 0};
};

struct ObjFile {
  std::vector<float> v, vt, vn;
  std::vector<ObjFace> faces;
  std::vector<std::pair<std::string, std::vector<size_t>>> objects;
};

size_t CvtIdx(int f, size_t count) { return f >= 0 ? size_t(f - 1) : count + size_t(f); }

bool ParseVertexRef(const std::string& tok, int& v, int& vt, int& vn) {
  v = vt = vn = 0;
  size_t a = tok.find('/');
  try {
    if (a == std::string::npos) {
      v = std::stoi(tok);
      return true;
    }
    v = std::stoi(tok.substr(0, a));
    size_t b = tok.find('/', a + 1);
    if (b == std::string::npos) {
      vt = std::stoi(tok.substr(a + 1));
      return true;
    }
    if (b > a + 1) vt = std::stoi(tok.substr(a + 1, b - a - 1));
    vn = std::stoi(tok.substr(b + 1));
    return true;
  } catch (...) {
    return false;
  }
}

ObjFile ReadObj(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw LoadError(RAD_ERR_IO, "cannot open file: " + path);
  ObjFile o;
  std::string line;
  int lineNum = 0;
  while (std::getline(f, line)) {
    lineNum++;
    std::istringstream ss(line);
    std::string cmd;
    if (!(ss >> cmd)) continue;
    if (cmd == "v" || cmd == "vn") {
      float x, y, z;
      if (!(ss >> x >> y >> z)) throw LoadError(RAD_ERR_IO, path + ": bad vertex at line " + std::to_string(lineNum));
      auto& dst = cmd == "v" ? o.v : o.vn;
      dst.insert(dst.end(), {x, y, z});
    } else if (cmd == "vt") {
      float x, y;
      if (!(ss >> x >> y)) throw LoadError(RAD_ERR_IO, path + ": bad uv at line " + std::to_string(lineNum));
      o.vt.insert(o.vt.end(), {x, y});
    } else if (cmd == "f") {
      std::vector<std::string> toks;
      std::string t;
      while (ss >> t) toks.push_back(t);
      if (toks.size() < 3) throw LoadError(RAD_ERR_IO, path + ": bad face at line " + std::to_string(lineNum));
      // the reference reads exactly three vertices per face (wavefront_obj_reader.cpp:205-248);
      // polygons are fanned so quads still load
      for (size_t k = 1; k + 1 < toks.size(); k++) {
        ObjFace face;
        const std::string* tri[3] = {&toks[0], &toks[k], &toks[k + 1]};
        for (int j = 0; j < 3; j++)
          if (!ParseVertexRef(*tri[j], face.v[j], face.vt[j], face.vn[j]))
            throw LoadError(RAD_ERR_IO, path + ": bad face at line " + std::to_string(lineNum));
        size_t index = o.faces.size();
        o.faces.push_back(face);
        if (!o.objects.empty()) o.objects.back().second.push_back(index);
      }
    } else if (cmd == "o" || cmd == "g") {
      std::string name;
      std::getline(ss, name);
      size_t b = name.find_first_not_of(" \t");
      name = b == std::string::npos ? "" : name.substr(b);
      while (!name.empty() && (name.back() == ' ' || name.back() == '\r')) name.pop_back();
      o.objects.push_back({name, {}});
    }
  }
  return o;
}

// WavefrontObjReader::ToModel: vertices deduplicated by (v, vn, vt) in order of first use.
Model ToModel(const ObjFile& o, const std::vector<size_t>* faceList) {
  Model m;
  struct Key {
    int p, n, t;
    bool operator==(const Key& k) const { return p == k.p && n == k.n && t == k.t; }
  };
  struct KeyHash {
    size_t operator()(const Key& k) const {
      return (size_t(uint32_t(k.p)) * 0x9E3779B97F4A7C15ull) ^ (size_t(uint32_t(k.n)) * 0xC2B2AE3D27D4EB4Full) ^
             (size_t(uint32_t(k.t)) * 0x165667B19E3779F9ull);
    }
  };
  std::unordered_map<Key, uint32_t, KeyHash> unique;
  size_t nv = o.v.size() / 3, nn = o.vn.size() / 3, nt = o.vt.size() / 2;
  size_t nf = faceList ? faceList->size() : o.faces.size();
  bool anyN = false, anyT = false;
  for (size_t fi = 0; fi < nf; fi++) {
    const ObjFace& f = o.faces[faceList ? (*faceList)[fi] : fi];
    for (int j = 0; j < 3; j++) {
      Key k{f.v[j], f.vn[j], f.vt[j]};
      auto it = unique.find(k);
      uint32_t index;
      if (it == unique.end()) {
        index = uint32_t(m.pos.size() / 3);
        unique.emplace(k, index);
        size_t p = CvtIdx(k.p, nv);
        if (p >= nv) throw LoadError(RAD_ERR_IO, "OBJ vertex index out of range");
        m.pos.insert(m.pos.end(), {o.v[3 * p], o.v[3 * p + 1], o.v[3 * p + 2]});
        if (k.n != 0) {
          size_t q = CvtIdx(k.n, nn);
          if (q >= nn) throw LoadError(RAD_ERR_IO, "OBJ normal index out of range");
          m.nrm.insert(m.nrm.end(), {o.vn[3 * q], o.vn[3 * q + 1], o.vn[3 * q + 2]});
          anyN = true;
        }
        if (k.t != 0) {
          size_t q = CvtIdx(k.t, nt);
          if (q >= nt) throw LoadError(RAD_ERR_IO, "OBJ uv index out of range");
          m.uv.insert(m.uv.end(), {o.vt[2 * q], o.vt[2 * q + 1]});
          anyT = true;
        }
      } else {
        index = it->second;
      }
      m.idx.push_back(index);
    }
  }
  size_t count = m.pos.size() / 3;
  if (anyN && m.nrm.size() != 3 * count) throw LoadError(RAD_ERR_IO, "OBJ mixes faces with and without normals");
  if (anyT && m.uv.size() != 2 * count) throw LoadError(RAD_ERR_IO, "OBJ mixes faces with and without uvs");
  return m;
}

struct Grid {
  int sx = 0, sy = 0, sz = 0, ch = 1;
  std::vector<float> data;
};

Grid ReadMitsubaVol(const std::string& path) {  // volume_reader.cpp:65-124
  std::ifstream f(path, std::ios::binary);
  if (!f) throw LoadError(RAD_ERR_IO, "cannot open file: " + path);
  char header[4];
  f.read(header, 4);
  if (!f || header[0] != 'V' || header[1] != 'O' || header[2] != 'L') throw LoadError(RAD_ERR_IO, "not a mitsuba vol file: " + path);
  if (header[3] != 3) throw LoadError(RAD_ERR_IO, "only version 3 vol files are supported");
  int32_t type, sx, sy, sz, ch;
  f.read((char*)&type, 4);
  if (type != 1) throw LoadError(RAD_ERR_IO, "vol file must hold float32 data");
  f.read((char*)&sx, 4);
  f.read((char*)&sy, 4);
  f.read((char*)&sz, 4);
  f.read((char*)&ch, 4);
  if (ch != 1 && ch != 3) throw LoadError(RAD_ERR_IO, "vol file must have 1 or 3 channels");
  float bbox[6];
  f.read((char*)bbox, 24);
  Grid g;
  g.sx = sx;
  g.sy = sy;
  g.sz = sz;
  g.ch = ch;
  g.data.resize(size_t(sx) * sy * sz * ch);
  f.read((char*)g.data.data(), g.data.size() * 4);
  if (!f) throw LoadError(RAD_ERR_IO, "truncated vol file: " + path);
  return g;
}

}  // namespace

// ------------------------------------------------------------------ loaded scene
struct rad_loaded_scene {
  rad_scene_desc desc{};
  rad_render_desc render{};
  std::string renderer_type, device;
  std::vector<rad_shape_desc> shapes;
  std::vector<rad_bsdf_desc> bsdfs;
  std::vector<rad_light_desc> lights;
  std::vector<rad_medium_desc> media;
  std::deque<Model> models;
  std::deque<Grid> grids;
  std::deque<std::vector<float>> images;  // linear texels of the image assets
  std::vector<rad_texture_desc> textures;
};

namespace {

struct Builder {
  rad_loaded_scene& out;
  std::string dir;
  std::map<std::string, const ObjFile*> objAssets;
  std::map<std::string, std::unique_ptr<ObjFile>> objStore;
  std::map<std::string, const Grid*> volAssets;
  std::set<std::string> vdbAssets;
  struct ImageAsset {
    const std::vector<float>* texels;
    uint32_t width, height;
    int channels;
    bool rgbOnly = false;  // image_exr: no single-channel view (image_exr.cpp:20-28)
  };
  std::map<std::string, ImageAsset> imageAssets;
  std::map<std::string, int> bsdfVars;
  rad_bsdf_desc* cur = nullptr;  // BSDF being built (its textured slots)

  std::string Resolve(const std::string& loc) const {
    if (!loc.empty() && loc[0] == '/') return loc;
    return dir.empty() ? loc : dir + "/" + loc;
  }

  void LoadAssets(const Value* assets) {
    if (!assets) return;
    if (!assets->is_array()) throw LoadError(RAD_ERR_INVALID_ARG, "assets must be an array");
    for (const Value& a : assets->arr) {
      std::string type = Str(a, "type"), name = Str(a, "name"), loc = Str(a, "location");
      if (type == "model_obj") {
        auto f = std::make_unique<ObjFile>(ReadObj(Resolve(loc)));
        objAssets[name] = f.get();
        objStore[name] = std::move(f);
      } else if (type == "volume_data_mitsuba_vol") {
        out.grids.push_back(ReadMitsubaVol(Resolve(loc)));
        volAssets[name] = &out.grids.back();
      } else if (type == "image_default") {  // asset/image_stb.cpp:12-100
        const int channel = IntOr(a, "channel", 3);
        const bool flip = BoolOr(a, "is_flip_y", true);
        const bool toLinear = BoolOr(a, "is_to_linear", channel == 3);
        rad::host::Image8 img;
        try {
          img = rad::host::ReadImage8(Resolve(loc), channel, flip);
        } catch (const std::runtime_error& e) {
          throw LoadError(RAD_ERR_IO, e.what());
        }
        std::vector<float> t(img.px.size());
        for (size_t i = 0; i < t.size(); i++) {
          float v = float(img.px[i]) / 255.0f;
          if (toLinear && channel == 3)  // Color24f::ToLinear (color.h:41-46)
            v = v <= 0.04045f ? v * (1.0f / 12.92f) : std::pow((v + 0.055f) * (1.0f / 1.055f), 2.4f);
          t[i] = v;
        }
        out.images.push_back(std::move(t));
        imageAssets[name] = ImageAsset{&out.images.back(), img.width, img.height, channel};
      } else if (type == "image_hdr") {  // asset/image_hdr.cpp:10-50
        rad::host::ImageF img;
        try {
          img = rad::host::ReadHdr(Resolve(loc), BoolOr(a, "is_flip_y", false));
        } catch (const std::runtime_error& e) {
          throw LoadError(RAD_ERR_IO, e.what());
        }
        out.images.push_back(std::move(img.px));
        imageAssets[name] = ImageAsset{&out.images.back(), img.width, img.height, 3};
      } else if (type == "image_exr") {  // asset/image_exr.cpp:31-97: RGB, no flip, no colour conversion
        rad::host::ImageF img;
        try {
          img = rad::host::ReadExr(Resolve(loc));
        } catch (const rad::host::ImageUnsupported& e) {
          throw LoadError(RAD_ERR_UNSUPPORTED, e.what());
        } catch (const std::runtime_error& e) {
          throw LoadError(RAD_ERR_IO, e.what());
        }
        out.images.push_back(std::move(img.px));
        imageAssets[name] = ImageAsset{&out.images.back(), img.width, img.height, 3, true};
      } else if (type == "volume_data_vdb") {  // OpenVDB grids are not carried: rejected where a medium uses them
        vdbAssets.insert(name);
      } else {
        throw LoadError(RAD_ERR_INVALID_ARG, "cannot find factory: " + type);
      }
    }
  }

  // A texture node (texture/bitmap.cpp:264-317, texture/chessboard.cpp:17-31);
  // returns 1 + its index in out.textures.
  int32_t MakeTexture(const Value& n, int channels) {
    std::string type = Str(n, "type");
    rad_texture_desc t{};
    t.channels = channels;
    const float ident[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    std::memcpy(t.to_uv, ident, sizeof(ident));
    if (type == "bitmap") {
      std::string asset = Str(n, "asset_name");
      auto it = imageAssets.find(asset);
      if (it == imageAssets.end()) throw LoadError(RAD_ERR_INVALID_ARG, "unknown image asset: " + asset);
      if (it->second.rgbOnly && channels != 3)
        throw LoadError(RAD_ERR_UNSUPPORTED, "exr image " + asset + " has no single-channel view");
      if (it->second.channels != channels)
        throw LoadError(RAD_ERR_INVALID_ARG, "loaded image " + asset + (channels == 3 ? " is mono" : " is rgb"));
      t.type = RAD_TEXTURE_BITMAP;
      t.width = it->second.width;
      t.height = it->second.height;
      t.texels = it->second.texels->data();
      std::string filter = StrOr(n, "filter", "bilinear");
      t.filter = filter == "nearest" ? RAD_FILTER_NEAREST
                 : filter == "trilinear" ? RAD_FILTER_TRILINEAR
                 : filter == "anisotropic" ? RAD_FILTER_ANISOTROPIC
                                           : RAD_FILTER_BILINEAR;  // unknown names fall back to bilinear
      t.wrap = StrOr(n, "wrap", "clamp") == "repeat" ? RAD_WRAP_REPEAT : RAD_WRAP_CLAMP;
      const bool mip = t.filter == RAD_FILTER_TRILINEAR || t.filter == RAD_FILTER_ANISOTROPIC;
      t.max_level = IntOr(n, "max_level", mip ? -1 : 0);
      if (!mip && t.max_level < 0) t.max_level = 0;
      t.anisotropic_level = FloatOr(n, "anisotropic_level", t.filter == RAD_FILTER_ANISOTROPIC ? 8.f : 0.f);
    } else if (type == "chessboard") {
      t.type = RAD_TEXTURE_CHESSBOARD;
      if (const Value* tu = n.find("to_uv")) {
        M4 m = AsTransform(*tu);
        std::memcpy(t.to_uv, m.m, sizeof(t.to_uv));
      }
      t.remap = BoolOr(n, "remap", false) ? 1 : 0;
      auto side = [&](const char* key, float def, float* color, int32_t& tex) {
        const Value* v = n.find(key);
        color[0] = color[1] = color[2] = def;
        if (!v) return;
        if (v->is_object()) {
          if (Str(*v, "type") != "bitmap")
            throw LoadError(RAD_ERR_UNSUPPORTED, std::string("chessboard ") + key + ": only constants and bitmaps nest");
          tex = MakeTexture(*v, channels);
        } else if (channels == 3) {
          Vec<3>(*v, color, key);
        } else {
          color[0] = Num(*v, key);
        }
      };
      side("color_a", 0.4f, t.color_a, t.tex_a);
      side("color_b", 0.2f, t.color_b, t.tex_b);
    } else {
      throw LoadError(RAD_ERR_INVALID_ARG, "cannot find factory: " + type);
    }
    out.textures.push_back(t);
    return int32_t(out.textures.size());
  }

  // ConfigNodeReadTexture (config_node_ext.h:17-45): a number / [r,g,b]
  // constant, or a texture object bound to the BSDF's slot.
  void Color(const Value& node, const char* key, float* out3, float def, int slot) {
    const Value* v = node.find(key);
    out3[0] = out3[1] = out3[2] = def;
    if (!v) return;
    if (v->is_object()) {
      cur->tex[slot] = MakeTexture(*v, 3);
      return;
    }
    Vec<3>(*v, out3, key);
  }
  // Area / skybox radiance (area.cpp:20-41, skybox.cpp:24-49): a constant,
  // or a texture (returns 1 + its index).
  int32_t EmitterColor(const Value& node, const char* key, float* out3) {
    const Value* v = node.find(key);
    out3[0] = out3[1] = out3[2] = 1.f;
    if (!v) return 0;
    if (v->is_object()) return MakeTexture(*v, 3);
    Vec<3>(*v, out3, key);
    return 0;
  }
  // Point intensity and a constant projection irradiance.
  void LightColor(const Value& node, const char* key, float* out3, float def) {
    const Value* v = node.find(key);
    out3[0] = out3[1] = out3[2] = def;
    if (!v) return;
    if (v->is_object()) throw LoadError(RAD_ERR_UNSUPPORTED, std::string("textured light '") + key + "' is not supported");
    Vec<3>(*v, out3, key);
  }
  float Mono(const Value& node, const char* key, float def, int slot) {
    const Value* v = node.find(key);
    if (!v) return def;
    if (v->is_object()) {
      cur->tex[slot] = MakeTexture(*v, 1);
      return def;
    }
    return Num(*v, key);
  }

  int MakeBsdf(const Value& n) {
    std::string type = Str(n, "type");
    rad_bsdf_desc b{};
    rad_bsdf_desc* outer = cur;
    cur = &b;
    b.sample_visible = 1;
    b.alpha_u = b.alpha_v = 0.1f;
    b.int_ior = 1.5046f;
    b.ext_ior = 1.000277f;
    b.specular = 0.5f;
    b.roughness = 0.5f;
    b.specular_sampling_rate = b.clearcoat_sampling_rate = b.diffuse_sampling_rate = 1.f;
    auto dist = [&](void) {
      std::string d = StrOr(n, "distribution", "ggx");
      b.distribution = d == "beckmann" ? RAD_MICROFACET_BECKMANN : RAD_MICROFACET_GGX;
      b.sample_visible = BoolOr(n, "sample_visible", true) ? 1 : 0;
      if (n.find("alpha")) {
        b.alpha_u = b.alpha_v = Mono(n, "alpha", 0.1f, RAD_TEX_ALPHA_U);
        b.tex[RAD_TEX_ALPHA_V] = b.tex[RAD_TEX_ALPHA_U];
      } else {
        b.alpha_u = Mono(n, "alpha_u", 0.1f, RAD_TEX_ALPHA_U);
        b.alpha_v = Mono(n, "alpha_v", 0.1f, RAD_TEX_ALPHA_V);
      }
    };
    if (type == "diffuse") {  // diffuse.cpp:40-44
      b.type = RAD_BSDF_DIFFUSE;
      Color(n, "reflectance", b.reflectance, 0.5f, RAD_TEX_REFLECTANCE);
    } else if (type == "rough_metal") {  // rough_metal.cpp:22-44
      b.type = RAD_BSDF_ROUGH_METAL;
      Color(n, "reflectance", b.reflectance, 1.f, RAD_TEX_REFLECTANCE);
      const float eta[3] = {0.1999899f, 0.9220823f, 1.0998751f}, k[3] = {3.9046354f, 2.4476316f, 2.1376500f};
      if (n.find("eta")) Color(n, "eta", b.eta, 0, RAD_TEX_ETA); else std::memcpy(b.eta, eta, sizeof(eta));
      if (n.find("k")) Color(n, "k", b.k, 0, RAD_TEX_K); else std::memcpy(b.k, k, sizeof(k));
      dist();
    } else if (type == "rough_glass") {  // rough_glass.cpp:22-48
      b.type = RAD_BSDF_ROUGH_GLASS;
      Color(n, "reflectance", b.reflectance, 1.f, RAD_TEX_REFLECTANCE);
      Color(n, "transmittance", b.transmittance, 1.f, RAD_TEX_TRANSMITTANCE);
      b.int_ior = FloatOr(n, "int_ior", 1.5046f);
      b.ext_ior = FloatOr(n, "ext_ior", 1.000277f);
      dist();
    } else if (type == "disney") {  // disney.cpp:62-110
      b.type = RAD_BSDF_DISNEY;
      Color(n, "base_color", b.base_color, 0.5f, RAD_TEX_BASE_COLOR);
      b.roughness = Mono(n, "roughness", 0.5f, RAD_TEX_ROUGHNESS);
      struct F {
        const char* key;
        uint32_t bit;
        float* dst;
        int slot;
      } fl[] = {{"anisotropic", RAD_DISNEY_HAS_ANISOTROPIC, &b.anisotropic, RAD_TEX_ANISOTROPIC},
                {"spec_trans", RAD_DISNEY_HAS_SPEC_TRANS, &b.spec_trans, RAD_TEX_SPEC_TRANS},
                {"sheen", RAD_DISNEY_HAS_SHEEN, &b.sheen, RAD_TEX_SHEEN},
                {"sheen_tint", RAD_DISNEY_HAS_SHEEN_TINT, &b.sheen_tint, RAD_TEX_SHEEN_TINT},
                {"flatness", RAD_DISNEY_HAS_FLATNESS, &b.flatness, RAD_TEX_FLATNESS},
                {"spec_tint", RAD_DISNEY_HAS_SPEC_TINT, &b.spec_tint, RAD_TEX_SPEC_TINT},
                {"metallic", RAD_DISNEY_HAS_METALLIC, &b.metallic, RAD_TEX_METALLIC},
                {"clearcoat", RAD_DISNEY_HAS_CLEARCOAT, &b.clearcoat, RAD_TEX_CLEARCOAT}};
      for (auto& e : fl) {
        const Value* v = n.find(e.key);
        // ConfigHasValue: present and (not a number, or a non-zero number)
        if (v && (!v->is_number() || v->num != 0)) b.disney_flags |= e.bit;
        *e.dst = Mono(n, e.key, 0.f, e.slot);
      }
      b.clearcoat_gloss = Mono(n, "clearcoat_gloss", 0.f, RAD_TEX_CLEARCOAT_GLOSS);
      b.specular_sampling_rate = FloatOr(n, "main_specular_sampling_rate", 1.f);
      b.clearcoat_sampling_rate = FloatOr(n, "clearcoat_sampling_rate", 1.f);
      b.diffuse_sampling_rate = FloatOr(n, "diffuse_reflectance_sampling_rate", 1.f);
      if (n.find("eta") && n.find("specular")) throw LoadError(RAD_ERR_INVALID_ARG, "eta cannot set with specular at same time");
      if (n.find("eta")) {
        b.disney_flags |= RAD_DISNEY_HAS_ETA;
        b.disney_eta = Num(*n.find("eta"), "eta");
      } else {
        b.specular = FloatOr(n, "specular", 0.5f);
      }
    } else if (type == "plastic" || type == "rough_plastic") {  // plastic.cpp:20-40, rough_plastic.cpp:20-55
      b.type = type == "plastic" ? RAD_BSDF_PLASTIC : RAD_BSDF_ROUGH_PLASTIC;
      b.int_ior = FloatOr(n, "int_ior", 1.49f);
      b.ext_ior = FloatOr(n, "ext_ior", 1.000277f);
      Color(n, "diffuse", b.diffuse, 0.5f, RAD_TEX_DIFFUSE);
      Color(n, "specular", b.specular_reflectance, 1.f, RAD_TEX_SPECULAR);
      b.nonlinear = BoolOr(n, "no_linear", false) ? 1 : 0;
      if (b.type == RAD_BSDF_ROUGH_PLASTIC) {
        std::string d = StrOr(n, "distribution", "ggx");
        b.distribution = d == "beckmann" ? RAD_MICROFACET_BECKMANN : RAD_MICROFACET_GGX;
        b.sample_visible = BoolOr(n, "sample_visible", true) ? 1 : 0;
        b.alpha_u = b.alpha_v = Mono(n, "alpha", 0.1f, RAD_TEX_ALPHA_U);
        b.tex[RAD_TEX_ALPHA_V] = b.tex[RAD_TEX_ALPHA_U];
      }
    } else if (type == "perfect_mirror") {  // perfect_mirror.cpp:16-22
      b.type = RAD_BSDF_PERFECT_MIRROR;
      Color(n, "reflectance", b.reflectance, 1.f, RAD_TEX_REFLECTANCE);
      Color(n, "eta", b.eta, 0.f, RAD_TEX_ETA);
      Color(n, "k", b.k, 1.f, RAD_TEX_K);
    } else if (type == "perfect_glass") {  // perfect_glass.cpp:16-23
      b.type = RAD_BSDF_PERFECT_GLASS;
      Color(n, "reflectance", b.reflectance, 1.f, RAD_TEX_REFLECTANCE);
      Color(n, "transmittance", b.transmittance, 1.f, RAD_TEX_TRANSMITTANCE);
      b.int_ior = FloatOr(n, "int_ior", 1.5046f);
      b.ext_ior = FloatOr(n, "ext_ior", 1.000277f);
    } else if (type == "two_side") {  // two_side.cpp:12-58
      const Value* inner = n.find("two_side");
      if (!inner)  // the out_side / in_side form reads uninitialised pointers in the reference (two_side.cpp:41-59)
        throw LoadError(RAD_ERR_UNSUPPORTED, "two_side: only the \"two_side\" form is supported");
      b.type = RAD_BSDF_TWO_SIDE;
      b.nested = MakeBsdf(*inner);
      cur = &b;
    } else if (type == "mask") {  // mask.cpp:8-17
      const Value* inner = n.find("nested");
      if (!inner) throw LoadError(RAD_ERR_INVALID_ARG, "mask: missing nested bsdf");
      b.type = RAD_BSDF_MASK;
      b.opacity = Mono(n, "opacity", 0.5f, RAD_TEX_OPACITY);
      b.nested = MakeBsdf(*inner);
      cur = &b;
    } else {
      throw LoadError(RAD_ERR_INVALID_ARG, "cannot find factory: " + type);
    }
    cur = outer;
    out.bsdfs.push_back(b);
    return int(out.bsdfs.size() - 1);
  }

  int MakeMedium(const Value& n, const M4& toWorld) {  // medium.cpp:8-17, homogeneous.cpp, heterogeneous.cpp
    std::string type = Str(n, "type");
    rad_medium_desc m{};
    const Value* ph = n.find("phase");
    if (!ph) throw LoadError(RAD_ERR_INVALID_ARG, "medium must have phase function");
    std::string pt = Str(*ph, "type");
    if (pt == "hg_phase") {
      m.phase = RAD_PHASE_HG;
      m.g = FloatOr(*ph, "g", 0.8f);
      if (m.g <= -1 || m.g >= 1) throw LoadError(RAD_ERR_INVALID_ARG, "param g must lie in the interval (-1, 1)");
    } else if (pt == "isotropic_phase") {
      m.phase = RAD_PHASE_ISOTROPIC;
    } else {
      throw LoadError(RAD_ERR_INVALID_ARG, "cannot find factory: " + pt);
    }
    std::memcpy(m.to_world, toWorld.m, sizeof(m.to_world));
    m.scale = FloatOr(n, "scale", 1.f);
    auto constVolume = [&](const char* key, float* dst, float def) {
      const Value* v = n.find(key);
      if (!v) {
        dst[0] = dst[1] = dst[2] = def;
        return false;
      }
      if (v->is_number() || v->is_array()) {
        Vec<3>(*v, dst, key);
        return false;
      }
      return true;
    };
    if (type == "homogeneous") {
      m.type = RAD_MEDIUM_HOMOGENEOUS;
      VecOr<3>(n, "sigma_t", m.sigma_t, {1, 1, 1});
      if (constVolume("albedo", m.albedo, 0.75f)) throw LoadError(RAD_ERR_UNSUPPORTED, "grid albedo is not supported");
    } else if (type == "heterogeneous") {
      m.type = RAD_MEDIUM_HETEROGENEOUS;
      if (constVolume("albedo", m.albedo, 0.75f)) throw LoadError(RAD_ERR_UNSUPPORTED, "grid albedo is not supported");
      if (constVolume("sigma_t", m.sigma_t, 1.f)) {
        const Value& v = *n.find("sigma_t");
        if (Str(v, "type") != "volume_grid") throw LoadError(RAD_ERR_INVALID_ARG, "cannot find factory: " + Str(v, "type"));
        std::string asset = Str(v, "asset_name");
        auto it = volAssets.find(asset);
        if (it == volAssets.end() && vdbAssets.count(asset))
          throw LoadError(RAD_ERR_UNSUPPORTED, "OpenVDB volume " + asset + " is not supported");
        if (it == volAssets.end()) throw LoadError(RAD_ERR_INVALID_ARG, "unknown asset: " + asset);
        const Grid* g = it->second;
        m.grid_size[0] = g->sx;
        m.grid_size[1] = g->sy;
        m.grid_size[2] = g->sz;
        m.grid_channels = g->ch;
        std::string wrap = StrOr(v, "wrap", "clamp");
        m.grid_wrap = wrap == "repeat" ? RAD_WRAP_REPEAT : RAD_WRAP_CLAMP;
        m.grid = g->data.data();
      }
    } else {
      throw LoadError(RAD_ERR_INVALID_ARG, "cannot find factory: " + type);
    }
    out.media.push_back(m);
    return int(out.media.size() - 1);
  }

  void Build(const Value& root) {
    // camera (perspective.cpp:15-29)
    const Value* cam = root.find("camera");
    if (!cam) throw LoadError(RAD_ERR_INVALID_ARG, "no child: camera");
    if (Str(*cam, "type") != "perspective") throw LoadError(RAD_ERR_INVALID_ARG, "cannot find factory: " + Str(*cam, "type"));
    rad_camera_desc& c = out.desc.camera;
    c.near_clip = FloatOr(*cam, "near", 0.001f);
    c.far_clip = FloatOr(*cam, "far", 100.f);
    c.fov = FloatOr(*cam, "fov", 30.f);
    float res[2];
    VecOr<2>(*cam, "resolution", res, {1280, 720});
    c.width = int(res[0]);
    c.height = int(res[1]);
    M4 ctw;
    if (const Value* tw = cam->find("to_world")) {
      ctw = AsTransform(*tw);
    } else {
      float o[3], t[3], u[3];
      VecOr<3>(*cam, "origin", o, {0, 0, -1});
      VecOr<3>(*cam, "target", t, {0, 0, 0});
      VecOr<3>(*cam, "up", u, {0, 1, 0});
      ctw = LookAtLeftHand(o, t, u);
    }
    std::memcpy(c.to_world, ctw.m, sizeof(c.to_world));
    // sampler (independent.cpp:15-20, sampler.cpp:5-7)
    const Value* sampler = root.find("sampler");
    out.render.spp = 16;
    out.render.seed = uint32_t(std::chrono::high_resolution_clock::now().time_since_epoch().count());
    if (sampler) {
      std::string st = StrOr(*sampler, "type", "independent");
      if (st != "independent") throw LoadError(RAD_ERR_INVALID_ARG, "cannot find factory: " + st);
      out.render.spp = uint32_t(IntOr(*sampler, "sample_count", 16));
      if (const Value* s = sampler->find("seed")) out.render.seed = uint32_t(int64_t(Num(*s, "seed")));
    }
    // global medium
    int globalMedium = -1;
    if (const Value* gm = root.find("global_medium")) globalMedium = MakeMedium(*gm, M4());
    // bsdf variables
    if (const Value* vars = root.find("bsdf_vars")) {
      for (const Value& v : vars->arr) {
        std::string name = Str(v, "name");
        const Value* b = v.find("bsdf");
        if (!b) throw LoadError(RAD_ERR_INVALID_ARG, "no child: bsdf");
        if (bsdfVars.count(name)) throw LoadError(RAD_ERR_INVALID_ARG, "duplicate bsdf var name: " + name);
        bsdfVars[name] = MakeBsdf(*b);
      }
    }
    // BFS over entities (build_context.cpp:137-288); child to_world = Child * Parent
    const Value* scene = root.find("scene");
    if (!scene || !scene->is_array()) throw LoadError(RAD_ERR_INVALID_ARG, "no child: scene");
    struct Ent {
      const Value* node;
      M4 toWorld;
    };
    std::deque<Ent> q;
    for (const Value& e : scene->arr) {
      M4 tw;
      if (const Value* t = e.find("to_world")) tw = AsTransform(*t);
      q.push_back({&e, tw});
    }
    while (!q.empty()) {
      Ent ent = q.front();
      q.pop_front();
      const Value& e = *ent.node;
      if (const Value* ch = e.find("children")) {
        for (const Value& c2 : ch->arr) {
          M4 tw = ent.toWorld;
          if (const Value* t = c2.find("to_world")) tw = AsTransform(*t) * ent.toWorld;
          q.push_back({&c2, tw});
        }
      }
      int shape = -1, light = -1, bsdf = -1, mIn = -1, mOut = -1;
      const Value* shapeNode = e.find("shape");
      const Value* lightNode = e.find("light");
      if (shapeNode) shape = MakeShape(*shapeNode, ent.toWorld);
      if (lightNode) {
        std::string lt = Str(*lightNode, "type");
        rad_light_desc l{};
        std::memcpy(l.to_world, ent.toWorld.m, sizeof(l.to_world));
        l.shape = -1;
        if (lt == "area") {
          l.type = RAD_LIGHT_AREA;
          l.tex = EmitterColor(*lightNode, "radiance", l.radiance);
          if (shape < 0) throw LoadError(RAD_ERR_INVALID_ARG, "area light entity without a shape");
          l.shape = shape;
        } else if (lt == "skybox") {
          l.type = RAD_LIGHT_SKYBOX;
          l.tex = EmitterColor(*lightNode, "radiance", l.radiance);
        } else if (lt == "point") {  // point.cpp:10-17: world position = to_world * position
          l.type = RAD_LIGHT_POINT;
          LightColor(*lightNode, "intensity", l.radiance, 1.f);
          float pos[3] = {0, 0, 0};
          VecOr<3>(*lightNode, "position", pos, {0, 0, 0});
          float w[3];
          for (int r = 0; r < 3; r++)
            w[r] = ent.toWorld.m[r * 4] * pos[0] + ent.toWorld.m[r * 4 + 1] * pos[1] + ent.toWorld.m[r * 4 + 2] * pos[2] +
                   ent.toWorld.m[r * 4 + 3];
          const float id[16] = {1, 0, 0, w[0], 0, 1, 0, w[1], 0, 0, 1, w[2], 0, 0, 0, 1};
          std::memcpy(l.to_world, id, sizeof(id));
        } else if (lt == "projection") {  // projection.cpp:14-36
          l.type = RAD_LIGHT_PROJECTION;
          float irr[3], sc[3];
          VecOr<3>(*lightNode, "scale", sc, {1, 1, 1});
          const Value* iv = lightNode->find("irradiance");
          if (iv && iv->is_object()) {  // an irradiance image: the device multiplies texels by the scale
            l.tex = MakeTexture(*iv, 3);
            for (int c = 0; c < 3; c++) l.radiance[c] = sc[c];
          } else {
            LightColor(*lightNode, "irradiance", irr, 1.f);
            for (int c = 0; c < 3; c++) l.radiance[c] = irr[c] * sc[c];
          }
          l.fov = FloatOr(*lightNode, "fov", 90.f);
        } else {
          throw LoadError(RAD_ERR_INVALID_ARG, "cannot find factory: " + lt);
        }
        out.lights.push_back(l);
        light = int(out.lights.size() - 1);
      }
      int bsdfRef = -1;
      if (const Value* b = e.find("bsdf")) {
        if (const Value* ref = b->find("ref")) {
          if (!ref->is_string()) throw LoadError(RAD_ERR_INVALID_ARG, "bsdf ref must be a string");
          auto it = bsdfVars.find(ref->str);
          if (it == bsdfVars.end()) throw LoadError(RAD_ERR_INVALID_ARG, "unknown bsdf var: " + ref->str);
          bsdfRef = it->second;
        } else {
          bsdf = MakeBsdf(*b);
        }
      }
      bool outUsesGlobal = true;
      if (const Value* mi = e.find("in_medium")) mIn = MakeMedium(*mi, ent.toWorld);
      if (const Value* mo = e.find("out_medium")) {
        if (const Value* ig = mo->find("is_global")) {
          if (!ig->is_bool()) throw LoadError(RAD_ERR_INVALID_ARG, "is_global must be bool");
          outUsesGlobal = ig->b;
        } else {
          mOut = MakeMedium(*mo, ent.toWorld);
        }
      }
      if (light >= 0 && shape >= 0 && bsdf < 0 && bsdfRef < 0) {  // build_context.cpp:233-236
        Value empty;
        empty.kind = Value::Object;
        Value t;
        t.kind = Value::String;
        t.str = "diffuse";
        empty.obj.push_back({"type", t});
        bsdf = MakeBsdf(empty);
      }
      if (shape >= 0) {
        rad_shape_desc& s = out.shapes[shape];
        s.light = light;
        s.bsdf = bsdf >= 0 ? bsdf : bsdfRef;
        s.medium_inside = mIn;
        s.medium_outside = mOut >= 0 ? mOut : (outUsesGlobal ? globalMedium : -1);
      }
    }
    // accel (default "embree"; any type maps to the device BVH)
    // renderer
    const Value* r = root.find("renderer");
    if (!r) throw LoadError(RAD_ERR_INVALID_ARG, "no child: renderer");
    out.renderer_type = Str(*r, "type");
    out.device = StrOr(*r, "device", "hip");
    out.render.integrator = out.renderer_type == "vol_path" ? RAD_INTEGRATOR_VOL_PATH : RAD_INTEGRATOR_PATH;
    out.render.max_depth = IntOr(*r, "max_depth", -1);
    out.render.rr_depth = IntOr(*r, "rr_depth", out.renderer_type == "vol_path" ? 5 : 3);
    out.render.guiding = BoolOr(*r, "guiding", false) ? 1 : 0;
    // finalize pointers
    out.desc.shape_count = uint32_t(out.shapes.size());
    out.desc.shapes = out.shapes.data();
    out.desc.bsdf_count = uint32_t(out.bsdfs.size());
    out.desc.bsdfs = out.bsdfs.data();
    out.desc.light_count = uint32_t(out.lights.size());
    out.desc.lights = out.lights.data();
    out.desc.medium_count = uint32_t(out.media.size());
    out.desc.media = out.media.data();
    out.desc.global_medium = globalMedium;
    out.desc.texture_count = uint32_t(out.textures.size());
    out.desc.textures = out.textures.data();
  }

  int MakeShape(const Value& n, const M4& toWorld) {
    std::string type = Str(n, "type");
    rad_shape_desc s{};
    s.bsdf = s.light = s.medium_inside = s.medium_outside = -1;
    std::memcpy(s.to_world, toWorld.m, sizeof(s.to_world));
    if (type == "mesh") {  // mesh.cpp:14-52
      s.type = RAD_SHAPE_MESH;
      std::string asset = Str(n, "asset_name");
      auto it = objAssets.find(asset);
      if (it == objAssets.end()) throw LoadError(RAD_ERR_INVALID_ARG, "unknown asset: " + asset);
      const std::vector<size_t>* faces = nullptr;
      if (const Value* sub = n.find("sub_model")) {
        if (!sub->is_string()) throw LoadError(RAD_ERR_INVALID_ARG, "sub_model must be a string");
        for (const auto& ob : it->second->objects)
          if (ob.first == sub->str) faces = &ob.second;
        if (!faces) throw LoadError(RAD_ERR_INVALID_ARG, "no model name: " + sub->str);
      }
      out.models.push_back(ToModel(*it->second, faces));
      const Model& m = out.models.back();
      s.vertex_count = uint32_t(m.pos.size() / 3);
      s.triangle_count = uint32_t(m.idx.size() / 3);
      s.positions = m.pos.data();
      s.normals = m.nrm.empty() ? nullptr : m.nrm.data();
      s.uvs = m.uv.empty() ? nullptr : m.uv.data();
      s.indices = m.idx.data();
    } else if (type == "sphere") {  // sphere.cpp:173-176
      s.type = RAD_SHAPE_SPHERE;
      VecOr<3>(n, "center", s.center, {0, 0, 0});
      s.radius = FloatOr(n, "radius", 1.f);
    } else if (type == "rectangle") {
      s.type = RAD_SHAPE_RECTANGLE;
      s.flip_normal = BoolOr(n, "is_flip_normal", false) ? 1 : 0;
    } else if (type == "cube") {
      s.type = RAD_SHAPE_CUBE;
    } else {
      throw LoadError(RAD_ERR_INVALID_ARG, "cannot find factory: " + type);
    }
    out.shapes.push_back(s);
    return int(out.shapes.size() - 1);
  }
};

}  // namespace

extern "C" {

rad_status rad_scene_load_json(const char* path, rad_loaded_scene** out) {
  if (!path || !out) {
    rad::SetLastError("NULL argument");
    return RAD_ERR_INVALID_ARG;
  }
  *out = nullptr;
  auto ls = std::make_unique<rad_loaded_scene>();
  try {
    std::ifstream f(path);
    if (!f) throw LoadError(RAD_ERR_IO, std::string("cannot open file: ") + path);
    std::stringstream ss;
    ss << f.rdbuf();
    Value root = rad::json::Parse(ss.str());
    Builder b{*ls};
    std::string p(path);
    size_t slash = p.find_last_of('/');
    b.dir = slash == std::string::npos ? "" : p.substr(0, slash);
    b.LoadAssets(root.find("assets"));
    b.Build(root);
  } catch (const LoadError& e) {
    rad::SetLastError(e.what());
    return e.code;
  } catch (const std::exception& e) {
    rad::SetLastError(e.what());
    return RAD_ERR_INVALID_ARG;
  }
  *out = ls.release();
  return RAD_OK;
}

const rad_scene_desc* rad_loaded_scene_desc(const rad_loaded_scene* s) { return s ? &s->desc : nullptr; }
const rad_render_desc* rad_loaded_scene_render(const rad_loaded_scene* s) { return s ? &s->render : nullptr; }
const char* rad_loaded_scene_renderer_type(const rad_loaded_scene* s) { return s ? s->renderer_type.c_str() : ""; }
const char* rad_loaded_scene_device(const rad_loaded_scene* s) { return s ? s->device.c_str() : ""; }
void rad_loaded_scene_free(rad_loaded_scene* s) { delete s; }

}  // extern "C"
