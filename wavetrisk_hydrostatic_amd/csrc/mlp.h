// Internal: what the free-surface solver uses of the learned correction.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/wavetrisk_gpu.h"

// bf16 input rows [max_batch][WTG_MLP_IN] and fp32 output rows [max_batch][WTG_MLP_OUT]
void wtg_mlp_buffers(wtg_mlp* m, uint16_t** x0, float** y, int* max_batch);
// forward pass of the first B rows of the input buffer into the output
// buffer, on `stream` (one launch; keeps no activations for a backward pass)
int wtg_mlp_infer_on(wtg_mlp* m, hipStream_t stream, int B);
const char* wtg_mlp_error_text(const wtg_mlp* m);
