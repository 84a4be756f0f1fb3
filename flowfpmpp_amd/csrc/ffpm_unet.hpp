// The learned preconditioner as seen by the solvers.
#pragma once

namespace ffpm {

struct UNet;

// One forward pass on features (FP32, 8 per cell, channels-last) and a source
// field (FP32); *out points at the network's FP32 output (valid until the
// next call on this network).
int unet_apply(UNet &net, int nx, int ny, int nz, const float *features, const float *source,
               const float **out);

}  // namespace ffpm
