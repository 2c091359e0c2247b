// Fixed per-device scratch for the entry points that allocate nothing (noalloc.hip).
#pragma once
#include <cstddef>

namespace rbc {
namespace noalloc {

// 32 MiB of doubles per device
constexpr size_t kScratchDoubles = (size_t)4 << 20;

// Device address of the calling thread's current device's scratch (looked up once per
// device with that device current); throws std::runtime_error when the lookup fails.
double* scratch_ptr();

}  // namespace noalloc
}  // namespace rbc
