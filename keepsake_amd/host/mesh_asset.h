// Mesh assets for the TOML scene path: MeshData / MeshAsset (geometry.h:22-84, mesh_asset.h:22-39),
// Wavefront OBJ loading with per-shape index remapping (mesh_asset.cpp:70-177, the tinyobj reader's shape
// and triangulation conventions), the binary mesh asset format (mesh_asset.cpp:180-241) and the
// [mesh_asset.*] table parser (create_mesh_asset, mesh_asset.cpp:243-264). Vertex arrays are stored
// without the reference's Embree padding floats; the binary format keeps them on disk.
#pragma once
#include "config.h"

#include <array>
#include <string>
#include <vector>

namespace ks {

struct MeshData {
    std::string name;
    std::vector<float> vertices;       // 3 per vertex
    std::vector<float> texcoords;      // 2 per vertex, or empty
    std::vector<float> vertex_normals; // 3 per vertex, or empty
    std::vector<uint32_t> indices;     // 3 per triangle
    bool twosided = true;
    bool use_smooth_normal = true;

    size_t vertex_count() const { return vertices.size() / 3; }
    size_t tri_count() const { return indices.size() / 3; }
    // MeshData::transform (geometry.cpp:15-32): points by the matrix, normals by its inverse transpose,
    // renormalised. m is row-major 4x4 (affine).
    void transform(const std::array<float, 16> &m);
};

struct MeshAsset {
    std::vector<MeshData> meshes;
};

// OBJ: one mesh per shape (a shape ends at an `o` / `g` statement that follows faces), polygons
// triangulated as a fan, (v, vt, vn) triples remapped to per-shape vertices in first-use order. Materials
// in MTL files are not read (load_materials must be false); texture coordinates are kept.
MeshAsset load_obj(const fs::path &path, bool twosided, bool use_smooth_normal);
// "i_am_a_binary_mesh_asset", int n_meshes, then per mesh: bool twosided, bool use_smooth_normal,
// int n_verts, int n_tris, bool has_texcoord, bool has_vertex_normal, float vertices[3 n + 1],
// uint32 indices[3 t], float texcoords[2 n + 2] (if any), float normals[3 n + 1] (if any).
MeshAsset load_mesh_binary(const fs::path &path);
void write_mesh_binary(const MeshAsset &asset, const fs::path &path);

// create_mesh_asset (mesh_asset.cpp:243-264): keys path, format ("obj" | "bin"), load_materials,
// twosided (false), use_smooth_normal (true), to_world.
MeshAsset create_mesh_asset(const ConfigArgs &args);

// convert_mesh_asset_task (mesh_asset.cpp:301-310): writes every asset named in `assets` as
// <task_dir>/<name>.bin.
void convert_mesh_asset(const ConfigArgs &args, const fs::path &task_dir, int task_id);

} // namespace ks
