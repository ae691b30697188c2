// Host-side internals of libkeepsake_amd shared between the HIP translation unit and the host-only
// C++ sources (BVH builder, procedural scenes). No device code here.
#pragma once
#include "../../include/keepsake_amd.h"

#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

namespace ksa {

// GPU BVH2 node: both children's boxes live in the parent so one 64-byte fetch tests two boxes.
//   f[0..3]   = c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y
//   f[4..7]   = c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y
//   f[8..11]  = c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z
//   child[2]  = child pointers (see kLeafBit), pad[2]; the host SBVH builder keeps the node's split axis
//               in pad[0] (read by the reference-order traversal count, not by the kernels)
// A child pointer without kLeafBit is the index of an interior node; with kLeafBit it is a leaf:
// bits 4..30 = first triangle reference, bits 0..3 = count - 1 (1..16 references).
struct alignas(16) GpuNode {
    float f[12];
    uint32_t child[2];
    uint32_t pad[2];
};
static_assert(sizeof(GpuNode) == 64, "GpuNode must be 64 bytes");

// GPU BVH4 node (the default): four children's boxes in SoA order, so one 128-byte line tests four
// boxes; child pointers as for GpuNode (interior: GpuNode4 index). Unused slots hold an empty box.
struct alignas(16) GpuNode4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(GpuNode4) == 128, "GpuNode4 must be 128 bytes");

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kMaxLeafRefs = 16;
constexpr uint32_t kMaxRefs = 1u << 27;

// One triangle reference in leaf order: the three vertices, v0.w carries the original triangle id.
struct alignas(16) GpuTri {
    float v0[3];
    uint32_t tri;
    float v1[4];
    float v2[4];
};
static_assert(sizeof(GpuTri) == 48, "GpuTri must be 48 bytes");

struct BvhResult {
    std::vector<GpuNode> nodes;   // BVH2
    std::vector<GpuNode4> nodes4; // the same tree collapsed to four children per node
    std::vector<GpuTri> tris;
    uint32_t max_depth = 0;       // leaf depth of the BVH2
    uint32_t max_depth4 = 0;      // leaf depth of the BVH4
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};

// SAH binned object + spatial split builder (restates cuda/sbvh.cu:301-440) followed by a flatten
// into GpuNode/GpuTri arrays.
BvhResult build_sbvh(const float *vertices, const uint32_t *indices, uint32_t n_tris, const ks_bvh_options &opt);
// Node fetches and triangle tests of the reference SBVH traversal order (cuda/sbvh.cu:512-603) over a
// host-built BVH2 for rays (n x 8: o, tmin, d, tmax); t_out (optional) gets each closest hit's distance
// (-1: miss) or the occlusion result (1 / -1).
void count_reference_traversal(const std::vector<GpuNode> &nodes, const std::vector<GpuTri> &tris, const float lo[3],
                               const float hi[3], const float *rays, int64_t n, bool any_hit, uint64_t &n_nodes,
                               uint64_t &n_tris, float *t_out);
// Collapses a flattened BVH2 into BVH4 nodes (leaves and references unchanged): each node repeatedly
// opens its interior child of largest surface area until it holds four children.
std::vector<GpuNode4> collapse_bvh4(const std::vector<GpuNode> &nodes, uint32_t &max_depth4);

// A BVH built on the device (ks_bvh_device.hip): node and triangle-record arrays in the layouts above,
// owned by the caller after a successful build (release with hipFree).
struct DeviceBvh {
    void *nodes = nullptr, *tris = nullptr;
    size_t node_bytes = 0, tri_bytes = 0;
    uint32_t n_nodes = 0, n_refs = 0, max_depth = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    double build_ms = 0.0;
    uint32_t rounds = 0; // clustering rounds (PLOC)
};
// verts: float4 per vertex, tri_idx: uint4 per triangle (vertex ids, material id in the low 24 bits of
// w), mat_alpha: per-material alpha-test flag or null, stream: the hipStream_t to build on. Returns 0 or
// -1 with `err` set.
int build_bvh_device(const void *verts, const void *tri_idx, uint32_t n_tris, const uint8_t *mat_alpha,
                     const ks_bvh_options &opt, int width, void *stream, DeviceBvh &out, std::string &err);

struct HostScene {
    std::vector<float> vertices, normals;
    std::vector<uint32_t> indices, tri_geom, geom_material;
    std::vector<ks_material_desc> materials;
    std::vector<ks_light_desc> lights;
    ks_camera_desc camera{};
    std::string name;
};

// Procedural scenes of the benchmark configurations; returns false for an unknown name.
bool make_procedural_scene(const std::string &name, uint64_t seed, HostScene &out);

void set_error(const std::string &msg);

// Runs the body of a C entry point so that no C++ exception crosses the ABI: std::bad_alloc becomes
// KS_ERR_OOM, any other exception KS_ERR_INVALID, each with its message in ks_last_error().
template <typename F> int guarded(F &&body) noexcept {
    auto note = [](const char *what, const char *detail) noexcept {
        try {
            set_error(std::string(what) + detail);
        } catch (...) {
        }
    };
    try {
        return body();
    } catch (const std::bad_alloc &) {
        note("host allocation failed", "");
        return KS_ERR_OOM;
    } catch (const std::exception &e) {
        note("internal error: ", e.what());
        return KS_ERR_INVALID;
    } catch (...) {
        note("internal error", "");
        return KS_ERR_INVALID;
    }
}

} // namespace ksa
