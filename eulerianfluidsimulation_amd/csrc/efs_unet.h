// Internal interface between the U-Net module and the learned preconditioner.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/efs.h"

// forward on `stream` from the BF16 input buffer (filled by the caller); *out16 receives the
// FP32 prediction (D, H, W), compact.  With `skip` (a device flag), every convolution returns at
// once when the flag is nonzero as it runs (the output is then left as it was).
int efs_unet_forward_on(efs_unet u, hipStream_t stream, int D, int H, int W, const float** out16, const int* skip = nullptr);
// BF16 input buffer (D, H, W, 16) for the given dims (allocated on demand)
int efs_unet_input_on(efs_unet u, int D, int H, int W, uint16_t** inb);
