// global_allocator.cuh — the reference's include path for the global API
// (`#include <gallatin/allocators/global_allocator.cuh>`, reference
// README.md:19-20), so that code written against it builds unchanged with
// hipcc. Everything lives in global_allocator.hpp.
#pragma once

#include "global_allocator.hpp"
