// JPEG decoding for glTF base-colour textures.
//
// The reference decodes images through tinygltf's bundled stb_image
// (libEyeRenderer3/MulticamScene.cpp:583-599), which reads baseline and
// progressive JPEG as well as PNG.  This decoder reads sequential
// Huffman-coded JPEG (ITU-T T.81: baseline SOF0 and extended SOF1, 8-bit
// samples): one to four components at any sampling factors from 1 to 4, one
// or several scans, restart intervals, 8- and 16-bit quantisation tables,
// JFIF YCbCr and Adobe RGB/YCbCr colour.  Progressive (SOF2), lossless and
// arithmetic-coded files are reported as unsupported.  Output is RGBA8.
//
// Reconstruction: dequantised coefficients go through a separable float IDCT
// (cosine table from double precision), +128 level shift, rounding and
// clamping.  Chroma subsampled by 2 along an axis (4:2:2, 4:4:0, 4:2:0) is
// upsampled with the triangle filter (weights 3/4 and 1/4 towards the nearer
// and farther chroma sample), other factors (4:1:1 ...) by replication, as
// stb_image and libjpeg do; upsampled samples are rounded to 8 bits, then
// converted with the JFIF equations.  Decoders with integer IDCTs (stb_image,
// libjpeg) may differ by a level or two per channel.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace eye {

// True when the bytes start with a JPEG start-of-image marker.
bool is_jpeg(const uint8_t* data, size_t size);

// Decodes a JPEG file held in memory into RGBA8 texels (alpha 255),
// row-major from the top row.  Returns false with a message on malformed or
// unsupported input.
bool jpeg_decode(const uint8_t* data, size_t size, int* width, int* height, std::vector<uint8_t>* rgba,
                 std::string* err);

}  // namespace eye
