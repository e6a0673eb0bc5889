// rd_hip_engine.cpp -- the MI355X engines inside the reference's own build.
//
// The file a maintainer adds to the reference (rdither) in place of its CUDA
// translation units src/rd_riemersma_cuda.cu, src/rd_blocks_cuda.cu,
// src/rd_approx_cuda.cu and the no-device stub src/rd_cuda_stub.cpp.  It is
// compiled against the REFERENCE's headers (include/rd_riemersma.h,
// rd_octree.h, rd_types.h) and this repository's C-ABI (include/stipple_amd.h),
// and linked with libstipple_amd.so beside the reference's own ColorTree
// (src/rd_octree.cpp) and CPU engines (src/rd_riemersma_cpu.cpp), which stay
// untouched.  It defines exactly what those four files define for the rest of
// the reference to call:
//   CudaAvailable, CudaMaxFrames          (src/rd_riemersma_cuda.cu:460, :523)
//   RiemersmaWalkCuda                     (src/rd_riemersma_cuda.cu:554)
//   RiemersmaBlocksCuda                   (src/rd_blocks_cuda.cu:1006)
//   CudaAllocPinned, CudaFreePinned       (src/rd_blocks_cuda.cu:1720, :1726)
//   RiemersmaApproxCuda, ApproxInverseFilter (src/rd_approx_cuda.cu:232, :386)
// The reference's ColorTree is built on the host, so each call uploads that
// tree's node arrays as the device search structure.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rd_octree.h"
#include "rd_riemersma.h"
#include "stipple_amd.h"

namespace rd {
namespace {
// The reference's ColorTree holds no device state: always upload.
sa_palette resident_palette(const Palette&, const ColorTree&) { return nullptr; }
}  // namespace

#include "rd_gpu_engines.inc"

}  // namespace rd
