// Image output of the small_pt task and the tonemap task: OpenEXR scanline files (write: FLOAT or HALF,
// ZIP or no compression; read: NONE / RLE / ZIPS / ZIP), Radiance .hdr input, PNG output, the
// composites RenderTarget2 saves (composite_and_save_to_exr / _png, render_target.cpp:127-247) and
// the tone mappers of tonemap.h (ACES fit, AgX with its looks, Khronos PBR Neutral). Files are written
// with this build's own encoders (zlib for deflate); tinyexr / stb are not used.
#pragma once
#include "config.h"

#include <array>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace ks {

// IEEE 754 binary16 from binary32, rounding to nearest even (overflow to infinity, NaN kept quiet)
uint16_t float_to_half(float f);
float half_to_float(uint16_t h);
// linear_to_srgb (maths.h:651-662)
float linear_to_srgb(float x);

enum class ExrCompression { None = 0, Zip = 3 };

// save_to_exr (image_util.cpp:101-190): `data` holds w*h*c interleaved floats (c in 1..4; channel
// names R, G, B, A), stored as HALF (round to nearest even) or FLOAT.
void write_exr(const fs::path &path, const float *data, int w, int h, int c, bool half,
               ExrCompression compression = ExrCompression::Zip);
// RGB floats of a scanline EXR (channels R, G, B by name, or Y replicated); PIZ, PXR24, B44 and DWA
// compressions and tiled or multi-part files are rejected with an error.
std::vector<float> read_exr(const fs::path &path, int &w, int &h);
// n_out = 1..4 floats per pixel: R, G, B (a single-channel image replicated into all four, as tinyexr's
// LoadEXR does) and A (1 when absent) -- the first n_out of them (load_from_exr, image_util.cpp:72-99).
std::vector<float> read_exr_channels(const fs::path &path, int &w, int &h, int n_out);
// RGB floats of a Radiance RGBE (.hdr) file (flat or run-length scanlines).
std::vector<float> read_hdr(const fs::path &path, int &w, int &h);
// 8-bit PNG, c = 1..4 channels, rows top to bottom.
void write_png(const fs::path &path, const uint8_t *data, int w, int h, int c);

// A decoded PNG as stb_image delivers it: palette images expanded to RGB (RGBA with tRNS), gray of 1, 2
// or 4 bits scaled to 8, tRNS colour keys turned into an alpha channel; 8- or 16-bit samples, rows top
// to bottom. Interlaced files are de-interlaced.
struct PngImage {
    int w = 0, h = 0, comp = 0, bits = 8;
    std::vector<uint8_t> data8;
    std::vector<uint16_t> data16;
    // stb_image's channel conversion (stbi__convert_format): luminance (77 r + 150 g + 29 b) >> 8,
    // opaque alpha where none was stored
    void to_channels(int c);
    // 16 -> 8 bits by the high byte (stbi__convert_16_to_8)
    void to_8bit();
};
PngImage decode_png(const uint8_t *bytes, size_t n, const std::string &name);
PngImage read_png(const fs::path &path);

// tonemap.h: ToneMapper::operator()(color3)
struct ToneMapper {
    virtual ~ToneMapper() = default;
    virtual std::array<float, 3> operator()(std::array<float, 3> c) const = 0;
};
// create_tone_mapper (tonemap.h:245-270): type "aces" | "agx" (look "none" | "golden" | "punchy") |
// "khronos_pbr_neutral", exposure (1).
std::unique_ptr<ToneMapper> create_tone_mapper(const ConfigArgs &args);

// The PNG composite of RenderTarget2 from its FP64 sums (W*H*3) and (hit, miss) weights (W*H*2):
// (sum + w_miss * backdrop) / (w_hit + w_miss), tone-mapped, sRGB-encoded, floor(255 c) clamped to
// [0, 255] (render_target.cpp:130-142). The EXR composite is ks_film_composite on the device.
std::vector<uint8_t> composite_png(const double *sums, const double *weights, size_t n, const float backdrop[3],
                                   const ToneMapper *tone_mapper);

// The tonemap task (tonemap.cpp:9-32): input_path (.exr or .hdr), output_path (.exr, FLOAT),
// tone_mapper table; paths relative to the asset root. Keys of this build: half (write HALF channels),
// png_path (also write the result sRGB-encoded as an 8-bit PNG), and tone_mapper may be left out to
// copy the image unchanged (an .hdr or EXR to a FLOAT or HALF EXR).
void tonemap_task(const ConfigArgs &args, const fs::path &task_dir, int task_id);

} // namespace ks
