// The small_pt task's resolved inputs: what the TOML keys of small_pt.cpp:180-255 turn into.
#pragma once
#include "../../include/ks/small_pt.h"
#include "config.h"
#include "image_io.h"

#include <memory>
#include <string>

namespace ks {

struct SmallPTSetup {
    SceneData scene;
    Camera camera;
    SmallPTInput input;
    std::string scene_name;
    bool guided = false;          // a [task.guide] table is present
    ks_guide_config guide_cfg{};  // its keys (zero: library default)
    std::shared_ptr<ToneMapper> tone_mapper; // [task.tone_mapper]: also write tone-mapped PNGs
    bool output_pfm = false;                 // also write PFM files (this build's key)
    int bvh_builder = KS_BVH_BUILD_HOST_SBVH; // bvh_builder = "sbvh" | "device" (this build's key)
};

// Parses one [[task]] table of type small_pt without touching the GPU.
SmallPTSetup resolve_small_pt(const ConfigArgs &args);
// The resolved inputs as one JSON object.
std::string describe(const SmallPTSetup &st);
// Device ordinal the small_pt task opens (ks_render -d).
void set_task_device(int ordinal);

} // namespace ks
