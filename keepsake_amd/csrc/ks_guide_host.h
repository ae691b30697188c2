// Host-side glue between the render translation unit (ks_api.hip) and the guide translation unit
// (ks_guide.hip): device accessors and the guide launchers the render loop calls.
#pragma once
#include "../../include/keepsake_amd.h"

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace ksr {
struct DevScene;
}

namespace ksg {

hipStream_t device_stream(ks_device *dev);
int device_ordinal(ks_device *dev);
int device_cu_count(ks_device *dev);

constexpr size_t kTrainSampleBytes = 64; // position + weight, normal, wo, wi (4 x float4)

ks_device *guide_device(const ks_guide *g);
float guide_fraction(const ks_guide *g);
int32_t guide_batch(const ks_guide *g);
int32_t guide_max_steps(const ks_guide *g);
int32_t guide_train_spp(const ks_guide *g);
int32_t guide_max_bounce(const ks_guide *g); // first bounce that is not guided (255: none)
// Mixtures (10 float4 per path slot) for the first *n_dev paths of the current wavefront buffer; paths
// whose bounce (state bits 16..23) is max_bounce or more, and misses, get none. The paths that get one
// are first compacted into `list` (n_max entries, their count in *n_list) so the inference's waves hold
// only real queries.
int launch_infer(ks_guide *g, const ksr::DevScene &sc, const uint32_t *n_dev, uint32_t n_max, const float4 *rd,
                 const float4 *hits, const uint32_t *state, int32_t max_bounce, float4 *mix, uint32_t *list,
                 uint32_t *n_list, hipStream_t s);
// One Adam step over n device-resident training samples samples[0], samples[stride], ... With c_scale > 0
// the sample weights are multiplied by c_scale and clamped at the guide's weight_clamp (render training:
// c_scale = 1 / mean weight of the pass); c_scale = 0 takes the weights as given.
int train_on_device(ks_guide *g, const void *samples, int n, int stride, float c_scale, double *loss_host,
                    hipStream_t s);
// Adam steps taken so far (0: the network has never been trained).
int64_t guide_steps(const ks_guide *g);
// Training samples from the path records of one pass: for each of n_items work items and each of its
// recorded vertices carrying `stamp`, weight c = luminance((L_final - L_before - NEE) / beta); records with
// c > 0 are appended to `out` (at most cap), counted in *count and their weights summed in *c_sum.
int launch_extract(const float4 *recs, const float4 *samp, uint32_t n_items, uint32_t stamp, void *out,
                   uint32_t *count, double *c_sum, uint32_t cap, hipStream_t s);

} // namespace ksg
