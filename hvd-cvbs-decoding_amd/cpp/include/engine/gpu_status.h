// engine/gpu_status.h -- failure reporting of the MI355X library behind the
// hvd-core API (no hvd-core counterpart).  hvd-core signals a failed decode
// with empty outputs; every entry point here does the same when the GPU is
// missing, a HIP call fails, or the configuration asks for work this library
// does not run (3-D temporal terms, diag_prior > 0), and records why.
#ifndef HVD_AMD_ENGINE_GPU_STATUS_H_
#define HVD_AMD_ENGINE_GPU_STATUS_H_

#include <string>

namespace hvd {

// Message of the last failure on this thread ("" if none since the last
// call to ClearGpuError).
std::string GpuLastError();
void ClearGpuError();

}  // namespace hvd

#endif  // HVD_AMD_ENGINE_GPU_STATUS_H_
