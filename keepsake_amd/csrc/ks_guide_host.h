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

float guide_fraction(const ks_guide *g);
int32_t guide_batch(const ks_guide *g);
int32_t guide_max_steps(const ks_guide *g);
// Mixtures (10 float4 per path slot) for the first *n_dev paths of the current wavefront buffer.
int launch_infer(ks_guide *g, const ksr::DevScene &sc, const uint32_t *n_dev, uint32_t n_max, const float4 *rd,
                 const float4 *hits, float4 *mix, hipStream_t s);
// One Adam step over n device-resident training samples.
int train_on_device(ks_guide *g, const void *samples, int n, double *loss_host, hipStream_t s);

} // namespace ksg
