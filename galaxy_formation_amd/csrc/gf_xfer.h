// Copies between pageable host memory (the numpy arrays of the host-array
// entry points) and the device.
#pragma once
#include "gf_common.h"

// Host -> device.  Returns once `src` may be reused; the device copy is
// ordered on stream s before any later work on s.
int gf_copy_h2d(void* dst_dev, const void* src_host, size_t bytes, hipStream_t s);
// Device -> host, after the work already on stream s.  Returns once `dst`
// holds the data.
int gf_copy_d2h(void* dst_host, const void* src_dev, size_t bytes, hipStream_t s);
