// Compound mesh assets from glTF 2.0 (.gltf with external or data-URI buffers, or binary .glb):
// CompoundMeshAsset::load_from_gltf and create_compound_mesh_asset (mesh_asset.cpp:311-1129). One
// prototype per glTF mesh, one MeshData per triangle primitive, one instance per node that carries a
// mesh, in depth-first order of the default scene with world = parent * local transforms. Materials
// convert to the principled BSDF from the constant glTF factors (baseColorFactor, metallicFactor,
// roughnessFactor, emissiveFactor times KHR_materials_emissive_strength, KHR_materials_ior,
// KHR_materials_specular, KHR_materials_transmission). The JSON reader is a small self-contained
// parser (tinygltf is not available).
#pragma once
#include "../../include/keepsake_amd.h"
#include "mesh_asset.h"
#include "texture_asset.h"

#include <array>
#include <string>
#include <vector>

namespace ks {

struct GltfLoadOptions {
    bool load_materials = false;
    // Texture fields, normal maps and non-opaque alpha modes have no place in the flat material record:
    // by default they are a load error; with this set the constant factors alone are kept.
    bool ignore_textures = false;
};

struct CompoundMeshAsset {
    std::vector<MeshAsset> prototypes;                    // per glTF mesh, one MeshData per primitive
    std::vector<std::vector<ks_material_desc>> materials; // per prototype and primitive (load_materials)
    // textured materials (load_materials): fields per prototype and primitive, used where `textured` is
    // set; their texture indices point into `textures`, one per glTF image
    std::vector<std::vector<ks_material_fields>> fields;
    std::vector<std::vector<uint8_t>> textured;
    std::vector<TextureData> textures;
    std::vector<std::pair<uint32_t, std::array<float, 16>>> instances; // (prototype, row-major to_world)
    bool has_materials = false;
};

CompoundMeshAsset load_gltf(const fs::path &path, const GltfLoadOptions &options);

// create_compound_mesh_asset (mesh_asset.cpp:1111-1129): keys path, format ("glb" | "gltf"; the file's
// own header decides which parser reads it), load_materials (required), bsdf_type ("principled_bsdf"),
// and this build's ignore_textures (false).
CompoundMeshAsset create_compound_mesh_asset(const ConfigArgs &args);

} // namespace ks
