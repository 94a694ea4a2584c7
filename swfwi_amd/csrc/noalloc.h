// Fixed per-device workspace of the allocation-free entry point
// (swfwi_forward_modeling_noalloc).
#pragma once
#include <cstddef>

#include "device.h"

namespace swfwi {

// 64 MiB per GPU: enough for a forward run on grids up to the 501x1501 one
// with nt=6000 (two fields, the [nt][ng] seismogram, the [nt][ns] sources).
constexpr size_t kNoAllocScratchFloats = size_t(16) << 20;

// Device address of `device`'s workspace (kNoAllocScratchFloats floats,
// 256-byte aligned); looked up once per device, with that device current.
float* noalloc_scratch(int device);

}  // namespace swfwi
