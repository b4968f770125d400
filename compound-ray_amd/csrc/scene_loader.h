// glTF 2.0 scene ingestion on the host (replaces loadScene / processGLTFNode /
// bufferViewFromGLTF of libEyeRenderer3/MulticamScene.cpp:85-736).  Meshes are
// flattened into world-space triangles with their shading colours; cameras
// (perspective, orthographic, panoramic, compound eye) keep their glTF extras.
#pragma once

#include <stdint.h>

#include <array>
#include <string>
#include <vector>

#include "eyecore.h"

namespace eye {

enum CameraKind { kPerspective = 0, kOrthographic = 1, kPanoramic = 2, kCompound = 3 };

struct CameraDesc {
  std::string name;
  CameraKind kind = kPerspective;
  Pose pose;                       // from the node transform
  float yfov_deg = 60.0f;
  float xmag = 1.0f, ymag = 1.0f;
  std::string eye_path;            // compound: the path actually opened
  std::string projection;          // compound: projection shader name
  std::vector<Ommatidium> ommatidia;
};

struct HitboxMesh {
  std::string name;
  std::vector<float> tris;         // world space, 9 floats per triangle
  float lo[3], hi[3];              // world AABB
};

// Base-colour texture: an RGBA8 image (texels first .. first + w*h - 1 of
// HostScene::texels, row 0 = top row) with its sampler's address modes and
// filter (MulticamScene.cpp:606-623, 801-834).
struct TextureDesc {
  int width = 0, height = 0;
  int wrap_s = kAddressWrap, wrap_t = kAddressWrap;
  int filter = kFilterLinear;
  uint32_t first = 0;
};

struct HostScene {
  std::vector<float> pos;          // 9 floats per triangle, world space
  // 9 floats per triangle: vertex colours, the base colour (x3), or for a
  // textured triangle its texture coordinates (u0 v0 u1 v1 u2 v2 0 0 0)
  std::vector<float> col;
  std::vector<uint32_t> mode;      // ColourMode per triangle (| texture index << 2)
  std::vector<uint32_t> texels;    // RGBA8, red in the low byte
  std::vector<TextureDesc> textures;
  std::vector<std::string> mesh_names;
  std::vector<uint32_t> mesh_first_tri;
  std::vector<std::array<float, 6>> mesh_bounds;  // world AABB per mesh primitive group (lo, hi)
  int sky_mode = 0;                // SkyMode
  std::string background_shader;   // "__miss__<name>"
  std::vector<CameraDesc> cameras;
  std::vector<HitboxMesh> hitboxes;
  std::vector<std::string> warnings;
  size_t tri_count() const { return mode.size(); }
};

// Parses `path`.  Never throws; returns false with `err` set on failure.
bool load_gltf(const std::string& path, HostScene* out, std::string* err);

// Reads a .eye file (8 space-separated floats per line: position, axis,
// acceptance angle, focal offset; data/eyes/eye-specification.txt).
bool read_eye_file(const std::string& path, std::vector<Ommatidium>* out, std::string* err);

// Axis-angle pose helpers (DataRecordCamera.h:66-87).
void pose_reset(Pose* p);
void pose_rotate_around(Pose* p, float angle, float ax, float ay, float az);
void pose_move(Pose* p, float x, float y, float z);

}  // namespace eye
