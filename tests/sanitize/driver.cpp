// AddressSanitizer + UndefinedBehaviorSanitizer driver for the host-side code
// that reads untrusted files or runs on the CPU only: the glTF/JSON reader,
// the PNG/DEFLATE and JPEG decoders, the .eye reader, the neighbour table and the CPU
// oracle.  Test infrastructure (built and run by tests/test_sanitizers.py);
// malformed inputs are expected to be rejected with an error, never to read
// or write out of bounds or hit undefined behaviour.
//
//   driver gltf FILE...        load_gltf on each file
//   driver png FILE...         png_decode on each file
//   driver jpeg FILE...        jpeg_decode on each file
//   driver eye FILE...         read_eye_file on each file
//   driver knn N K SEED        nearest_ommatidia on N seeded ommatidia
//   driver render FILE SPP     load_gltf, then the oracle's BVH build and a
//                              render of the first compound eye (or 64
//                              synthetic ommatidia) at SPP samples
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <string>
#include <vector>

#include "jpeg.h"
#include "knn.h"
#include "png.h"
#include "scene_loader.h"

extern "C" {
void* oracle_scene_create(const float* pos, const float* col, const uint32_t* mode, int64_t n, int sky,
                          int leaf_max);
void oracle_scene_free(void* h);
void oracle_scene_set_textures(void* h, const int32_t* desc, int64_t ntex, const uint32_t* texels, int64_t ntexels);
void oracle_render(void* h, const float* omms, int64_t n_omm, const uint32_t* counts, int counts_per_unit, int spp,
                   const float* poses, int64_t n_pose, uint32_t seed, uint32_t eye, uint32_t frame,
                   uint32_t sample_base, float* samples, float* col_gpu, float* col_seq, double* col_f64,
                   uint64_t* visits, int threads);
void oracle_trace(void* h, const float* rays, int64_t n, int mode, int32_t* ids, float* tbb, uint64_t* visits,
                  int threads);
void oracle_generate_rays(const float* omms, int64_t n_omm, const uint32_t* counts, int spp, const float* poses,
                          int64_t n_pose, uint32_t seed, uint32_t eye, uint32_t frame, float* out);
}

namespace {

bool read_all(const char* path, std::vector<uint8_t>* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  out->assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return true;
}

int cmd_gltf(int argc, char** argv) {
  for (int i = 0; i < argc; ++i) {
    eye::HostScene scene;
    std::string err;
    const bool ok = eye::load_gltf(argv[i], &scene, &err);
    std::printf("%s: %s (%zu triangles)\n", argv[i], ok ? "ok" : err.c_str(), scene.tri_count());
  }
  return 0;
}

int cmd_png(int argc, char** argv) {
  for (int i = 0; i < argc; ++i) {
    std::vector<uint8_t> bytes, rgba;
    if (!read_all(argv[i], &bytes)) return 2;
    int w = 0, h = 0;
    std::string err;
    const bool ok = eye::png_decode(bytes.data(), bytes.size(), &w, &h, &rgba, &err);
    unsigned sum = 0;
    for (uint8_t v : rgba) sum += v;   // touches every decoded byte
    std::printf("%s: %s %dx%d sum %u\n", argv[i], ok ? "ok" : err.c_str(), w, h, sum);
  }
  return 0;
}

int cmd_jpeg(int argc, char** argv) {
  for (int i = 0; i < argc; ++i) {
    std::vector<uint8_t> bytes, rgba;
    if (!read_all(argv[i], &bytes)) return 2;
    int w = 0, h = 0;
    std::string err;
    const bool ok = eye::jpeg_decode(bytes.data(), bytes.size(), &w, &h, &rgba, &err);
    unsigned sum = 0;
    for (uint8_t v : rgba) sum += v;
    std::printf("%s: %s %dx%d sum %u\n", argv[i], ok ? "ok" : err.c_str(), w, h, sum);
  }
  return 0;
}

int cmd_eye(int argc, char** argv) {
  for (int i = 0; i < argc; ++i) {
    std::vector<eye::Ommatidium> omm;
    std::string err;
    const bool ok = eye::read_eye_file(argv[i], &omm, &err);
    std::printf("%s: %s (%zu ommatidia)\n", argv[i], ok ? "ok" : err.c_str(), omm.size());
  }
  return 0;
}

std::vector<float> random_ommatidia(size_t n, uint32_t seed) {
  std::mt19937 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  std::vector<float> o(8 * n);
  for (size_t i = 0; i < n; ++i) {
    float* r = &o[8 * i];
    r[3] = nd(rng); r[4] = nd(rng); r[5] = nd(rng);
    r[0] = 0.001f * r[3]; r[1] = 0.001f * r[4]; r[2] = 0.001f * r[5];
    r[6] = 0.04f;
    r[7] = 0.0f;
  }
  // a few exact duplicates and a zero axis
  if (n > 8) {
    std::memcpy(&o[8 * (n - 1)], &o[0], 8 * sizeof(float));
    std::memcpy(&o[8 * (n - 2)], &o[0], 8 * sizeof(float));
    o[8 * 3 + 3] = o[8 * 3 + 4] = o[8 * 3 + 5] = 0.0f;
  }
  return o;
}

int cmd_knn(int argc, char** argv) {
  if (argc < 3) return 2;
  const size_t n = std::strtoul(argv[0], nullptr, 10);
  const int k = std::atoi(argv[1]);
  const uint32_t seed = (uint32_t)std::strtoul(argv[2], nullptr, 10);
  std::vector<float> o = random_ommatidia(n, seed);
  std::vector<uint32_t> out(n * (size_t)k + 1);
  eye::nearest_ommatidia(o.data(), n, k, eye::kNeighbourPosScale, out.data());
  uint64_t sum = 0;
  for (size_t i = 0; i < n * (size_t)k; ++i) {
    if (out[i] >= n) return 3;
    sum += out[i];
  }
  std::printf("knn %zu x %d: sum %llu\n", n, k, (unsigned long long)sum);
  return 0;
}

int cmd_render(int argc, char** argv) {
  if (argc < 2) return 2;
  eye::HostScene scene;
  std::string err;
  if (!eye::load_gltf(argv[0], &scene, &err)) {
    std::printf("%s: %s\n", argv[0], err.c_str());
    return 0;
  }
  const int spp = std::atoi(argv[1]);
  const int64_t n = (int64_t)scene.tri_count();
  void* h = oracle_scene_create(scene.pos.data(), scene.col.data(), scene.mode.data(), n, scene.sky_mode, 4);
  if (!scene.textures.empty()) {
    std::vector<int32_t> desc;
    for (const eye::TextureDesc& t : scene.textures) {
      desc.push_back(t.width);
      desc.push_back(t.height);
      desc.push_back(t.wrap_s);
      desc.push_back(t.wrap_t);
      desc.push_back(t.filter);
      desc.push_back((int32_t)t.first);
    }
    oracle_scene_set_textures(h, desc.data(), (int64_t)scene.textures.size(), scene.texels.data(),
                              (int64_t)scene.texels.size());
  }
  std::vector<float> omm;
  std::vector<float> pose = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (const eye::CameraDesc& c : scene.cameras)
    if (c.kind == eye::kCompound && !c.ommatidia.empty()) {
      omm.assign(reinterpret_cast<const float*>(c.ommatidia.data()),
                 reinterpret_cast<const float*>(c.ommatidia.data()) + 8 * c.ommatidia.size());
      std::memcpy(pose.data(), &c.pose, sizeof(float) * 12);
      break;
    }
  if (omm.empty()) omm = random_ommatidia(64, 5);
  const int64_t m = (int64_t)(omm.size() / 8);
  const size_t total = (size_t)m * (size_t)spp;
  std::vector<float> samples(4 * total), gpu(3 * (size_t)m), seq(3 * (size_t)m);
  std::vector<double> f64(3 * (size_t)m);
  std::vector<uint64_t> visits(2);
  oracle_render(h, omm.data(), m, nullptr, 0, spp, pose.data(), 1, 42, 0, 0, 0, samples.data(), gpu.data(),
                seq.data(), f64.data(), visits.data(), 2);
  // uneven counts, and the brute-force trace of the same rays
  std::vector<uint32_t> counts((size_t)m);
  for (int64_t i = 0; i < m; ++i) counts[(size_t)i] = (uint32_t)(i % 5);
  oracle_render(h, omm.data(), m, counts.data(), 0, 1, pose.data(), 1, 42, 0, 1, 0, nullptr, gpu.data(), seq.data(),
                nullptr, visits.data(), 2);
  std::vector<float> rays(8 * total);
  oracle_generate_rays(omm.data(), m, nullptr, spp, pose.data(), 1, 42, 0, 0, rays.data());
  const int64_t nr = (int64_t)std::min<size_t>(total, 2000);
  std::vector<int32_t> ids((size_t)nr), ids_bf((size_t)nr);
  std::vector<float> tbb(3 * (size_t)nr);
  oracle_trace(h, rays.data(), nr, 1, ids.data(), tbb.data(), nullptr, 2);
  oracle_trace(h, rays.data(), nr, 0, ids_bf.data(), tbb.data(), nullptr, 2);
  oracle_scene_free(h);
  std::printf("%s: rendered %lld ommatidia x %d samples; BVH and brute force agree: %s\n", argv[0], (long long)m, spp,
              ids == ids_bf ? "yes" : "no");
  return ids == ids_bf ? 0 : 4;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string cmd = argv[1];
  if (cmd == "gltf") return cmd_gltf(argc - 2, argv + 2);
  if (cmd == "png") return cmd_png(argc - 2, argv + 2);
  if (cmd == "jpeg") return cmd_jpeg(argc - 2, argv + 2);
  if (cmd == "eye") return cmd_eye(argc - 2, argv + 2);
  if (cmd == "knn") return cmd_knn(argc - 2, argv + 2);
  if (cmd == "render") return cmd_render(argc - 2, argv + 2);
  return 2;
}
