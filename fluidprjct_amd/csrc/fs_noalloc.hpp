// fs_noalloc.hpp -- the fixed device workspace of the allocation-free entry point.
#pragma once

#include <hip/hip_runtime.h>

namespace fsk {

// Device address of the current device's copy of the workspace
// (FS_NOALLOC_SCRATCH_BYTES bytes, 256-byte aligned), looked up on the first
// call on each device; nullptr and *err on failure.  Allocates nothing.
void* noalloc_scratch(int device, hipError_t* err);

}  // namespace fsk
