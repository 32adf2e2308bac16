// server_kernels.hip -- batched device-side copies of the serving loop.
//
// Admission and retirement move many small pieces (one request's tables
// into its arena ranges, final positions out of them).  The host packs them
// into one staging buffer and one descriptor list; k_copy_desc performs all
// copies in one launch (one workgroup per descriptor) instead of one API
// call per piece.
#include <hip/hip_runtime.h>

#include "smr_kernels.h"

namespace smr {

__global__ __launch_bounds__(256) void k_copy_desc(const CopyDesc *__restrict__ d) {
    const CopyDesc c = d[blockIdx.x];
    if (((c.src | c.dst | c.bytes) & 15ull) == 0) {
        const uint4 *s = reinterpret_cast<const uint4 *>(c.src);
        uint4 *t = reinterpret_cast<uint4 *>(c.dst);
        const size_t n = c.bytes >> 4;
        for (size_t i = threadIdx.x; i < n; i += blockDim.x) t[i] = s[i];
    } else {
        // every piece is a whole number of 16-bit words, 2-byte aligned
        const unsigned short *s = reinterpret_cast<const unsigned short *>(c.src);
        unsigned short *t = reinterpret_cast<unsigned short *>(c.dst);
        const size_t n = c.bytes >> 1;
        for (size_t i = threadIdx.x; i < n; i += blockDim.x) t[i] = s[i];
    }
}

hipError_t launch_copy_desc(const CopyDesc *d, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_desc, dim3(n), dim3(256), 0, st, d);
    return hipGetLastError();
}

}  // namespace smr
