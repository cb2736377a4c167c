// gallatin.cuh — the reference's include path for Gallatin<segment_size,
// min_size, max_size> (reference include/gallatin/allocators/gallatin.cuh),
// so that code written against it builds unchanged with hipcc. Everything
// lives in gallatin.hpp.
#pragma once

#include "gallatin.hpp"
