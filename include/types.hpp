// types.hpp -- global constants, the COO execution mode and the error check
// of the HBDIA API.  Mirrors include/types.hpp of the reference (:8-21); the
// values are kept so reference call sites compile unchanged.
#ifndef TYPES_HPP
#define TYPES_HPP

#include <stdexcept>
#include <string>

#include "hbdia_hip.h"

using DataType = double;

// Where the COO fallback runs (reference include/types.hpp:8-11).  GPU_COO
// uploads the fallback to the device at prepareForGPU time.
enum class ExecutionMode {
  CPU_COO,
  GPU_COO
};

#define DEFAULT_BLOCK_WIDTH 32
#define DEFAULT_THRESHOLD 1
#define MAX_COO_ENTRIES 1000000000
#define THREADS_PER_BLOCK_SPMV 256
#define THREADS_PER_BLOCK_VECTOR_ADD 256

#define RAND_SEED 42
#define NUM_BENCH_ITER 100

namespace hbdia {
// Raised for any failing device call.  The reference's CHECK_CUDA prints and
// calls exit(1) (include/types.hpp:37-44); a library never exits, so this
// build throws instead.
class DeviceError : public std::runtime_error {
 public:
  DeviceError(const std::string& what, int code) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};
}  // namespace hbdia

#define CHECK_HBDIA(call)                                                                  \
  do {                                                                                     \
    int hbdia_rc_ = (call);                                                                \
    if (hbdia_rc_ != HBDIA_SUCCESS)                                                        \
      throw ::hbdia::DeviceError(std::string(__FILE__) + ":" + std::to_string(__LINE__) +  \
                                     " - " #call " - " + hbdia_last_error(),               \
                                 hbdia_rc_);                                               \
  } while (0)

#endif  // TYPES_HPP
