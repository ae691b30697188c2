// Texture assets of the TOML scene path: [texture.*] tables (create_texture, texture.cpp:1068-1084),
// images read as the reference's create_texture_from_image does (texture.cpp:922-948: .exr / .hdr as
// f32, other images as u8 -- PNG here), the serialized texture format with its LZ4 blocks
// (texture.cpp:950-1066), the convert_texture task (texture.cpp:1123-1148) and shader-field tables
// (create_shader_field_color, shader_field.h:88-158; create_texture_sampler, texture.cpp:1086-1121).
#pragma once
#include "../../include/keepsake_amd.h"
#include "config.h"

#include <map>
#include <string>
#include <vector>

namespace ks {

struct TextureData {
    std::vector<uint8_t> bytes; // levels back to back, texels of `channels` values of `data_type`
    int width = 0, height = 0, channels = 0;
    int data_type = KS_TEX_U8, color_space = KS_COLORSPACE_LINEAR;
    int levels = 1;        // levels present in bytes
    int build_mipmaps = 0; // with levels == 1: the library builds the chain
    ks_texture_desc desc() const;
};

// create_texture_from_image: `channels` channels of an .exr / .hdr (f32) or PNG (u8) image.
TextureData texture_from_image(int channels, bool build_mipmaps, int color_space, const fs::path &path);
// create_texture_from_serialized / write_texture_to_serialized: "i_am_a_serialized_texture", width,
// height, channels, data type (int32), levels, total bytes (size_t), LZ4 block count and sizes, blocks.
// Read textures are linear, as in the reference.
TextureData read_serialized_texture(const fs::path &path);
void write_serialized_texture(const TextureData &t, const fs::path &path); // writes the full chain when asked
// create_texture: path, serialized (false), channels, build_mipmaps, colorspace ("Linear" | "sRGB").
TextureData create_texture(const ConfigArgs &args);
void convert_texture(const ConfigArgs &args, const fs::path &task_dir, int task_id);

// LZ4 block format (the LZ4 project's published block specification): greedy compression, bounded
// decompression that returns the decoded size or throws on malformed input.
std::vector<uint8_t> lz4_compress(const uint8_t *src, size_t n);
size_t lz4_decompress(const uint8_t *src, size_t n, uint8_t *dst, size_t cap);

// Textures of one scene by asset name, in first-use order.
struct TextureRegistry {
    std::vector<TextureData> textures;
    std::map<std::string, int> index;
    int get(const ConfigArgs &args, const std::string &name); // loads [texture.*] `name` once
};

// A shader field table: { type = "constant", value = ... } or { type = "texture", map = "texture.X",
// sampler = { type = "nearest" | "linear" | "cubic" | "ewa", wrap_mode_u, wrap_mode_v, kernel,
// anisotropy }, flip_v (true), swizzle, uv_scale, uv_offset, scale } with `channels` components.
ks_field_desc load_shader_field(const ConfigArgs &field, int channels, TextureRegistry &reg);

} // namespace ks
