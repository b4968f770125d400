// PNG decoding for glTF base-colour textures.
//
// The reference decodes images through tinygltf's bundled stb_image and
// asserts four 8-bit (or 16-bit) components per texel
// (libEyeRenderer3/MulticamScene.cpp:583-599).  This decoder covers what PNG
// files can hold: colour types 0/2/3/4/6, bit depths 1-16, palettes with tRNS
// transparency, all five row filters and Adam7 interlacing, with its own
// DEFLATE decoder (RFC 1950/1951).  Output is always RGBA with 8 bits per
// channel; 16-bit samples are rounded to 8 bits (v * 255 / 65535).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace eye {

// Inflates a zlib stream (2-byte header, DEFLATE blocks, Adler-32 trailer).
// Returns false with a message on malformed input or a checksum mismatch.
bool zlib_inflate(const uint8_t* data, size_t size, std::vector<uint8_t>* out, std::string* err);

// Decodes a PNG file held in memory into RGBA8 texels, row-major from the
// top row.
bool png_decode(const uint8_t* data, size_t size, int* width, int* height, std::vector<uint8_t>* rgba,
                std::string* err);

}  // namespace eye
