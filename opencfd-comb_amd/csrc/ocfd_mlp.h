// ocfd_mlp.h -- internal view of the MLP for the other translation units of
// the library (host side: device pointers and padded shapes).
#pragma once
#include <hip/hip_runtime.h>

#define MLP_MAX_LAYERS 8

struct MlpDevice {
  int n_layers;
  int din[MLP_MAX_LAYERS], dout[MLP_MAX_LAYERS];    // true widths
  int dinp[MLP_MAX_LAYERS], doutp[MLP_MAX_LAYERS];  // padded widths (storage)
  const __bf16* W[MLP_MAX_LAYERS];                  // [doutp][dinp] row-major, zero padded
  const float* b[MLP_MAX_LAYERS];                   // [doutp]
};

// NULL until ocfd_mlp_create succeeded
const MlpDevice* ocfd_mlp_device_view();
// changes whenever the bf16 weights change (creation, set_params, train_step)
unsigned long long ocfd_mlp_weights_version();
hipStream_t ocfd_mlp_stream_handle();
