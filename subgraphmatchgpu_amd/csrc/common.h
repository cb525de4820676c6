// common.h -- shared scalar types, limits and error plumbing of the engine.
//
// Vertex and edge id widths follow the reference (src/Meta.h:8-9): vertex ids
// are uint32, CSR offsets are uint64.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

namespace sme {

using uintV = uint32_t;  // vertex id (Meta.h:8)
using uintE = uint64_t;  // edge offset (Meta.h:9)

constexpr int kMaxQueryVertices = 6;  // Meta.h:30
constexpr int kMaxOps = 16;
constexpr int kNumQueries = 14;  // QueryType Q0..Q13, Meta.h:104-124

// Status codes returned through the C ABI (include/sme.h).
enum Status : int {
  kOk = 0,
  kErrInvalid = 1,
  kErrOom = 2,
  kErrHip = 3,
  kErrIo = 4,
  kErrUnsupported = 5,
};

// Error carrying a status code; the C ABI catches it and stores the message
// in the context (sme_last_error) instead of exiting, unlike the reference's
// CUDA_ERROR -> exit(-1) (src/gpu/utils/GPUUtil.cuh:9-17).
struct Error {
  Status status;
  std::string msg;
};

[[noreturn]] void Fail(Status s, const std::string& msg);

}  // namespace sme
