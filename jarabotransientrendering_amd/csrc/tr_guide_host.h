// Internal interface between the guided render loop (tr_kernels.hip) and the guiding network's
// host object (tr_guide.hip). Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tr_api.h"

namespace trg {

// What the shade and film kernels need to use and train the network during a render.
struct GuideView {
    const uint4* frag;     // forward weight fragments (bf16)
    const float* bias;     // padded biases, 64 per layer
    float4* rec;           // training records [train_paths * max_levels][4]
    uint32_t* n_rec;       // records written per training path
    float* contrib;        // luminance of visible NEE contributions [train_paths * max_levels]
    uint32_t train_paths;  // training paths per wave (slot % stride == 0)
    uint32_t stride;       // slot stride between training paths
    uint32_t max_levels;   // records per training path
    float alpha;           // probability of sampling the BSDF / phase function
    float3 bb_lo, bb_inv;  // scene box used to normalise positions
};

// Allocates per-render buffers and fills the view for a wave of `wave` paths.
int guide_begin_render(tr_guide g, uint32_t wave, uint32_t max_levels, const float bb_lo[3], const float bb_hi[3],
                       hipStream_t stream, GuideView* view);
// Before a wave: clears the record counts and contribution sums.
int guide_wave_begin(tr_guide g, hipStream_t stream);
// After a wave: builds the training batch from the wave's records and runs the Adam steps.
// `ev_begin`/`ev_end`, if not null, bracket the work for kernel timing.
int guide_wave_end(tr_guide g, hipStream_t stream);

}  // namespace trg
