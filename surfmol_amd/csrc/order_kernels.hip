// order_kernels.hip -- start order of a batched relaxation, on the device.
//
//  k_reset_state  FIRE state of every system at the start of a relaxation
//  k_rank_keys    sort key of every work-list position: its launch slice,
//                 then atom count (descending), then FP64 energy per atom
//                 (descending) from the state's energy terms
//  rank_work_list the keys, then a stable radix sort (hipCUB) that writes
//                 the work list the relax launches read
//
// Everything is issued on the caller's stream without a host round trip, so
// a relaxation's ranking, launches and final energies run back to back.
#include <hipcub/hipcub.hpp>

#include "smr_internal.h"

namespace smr {

__global__ void k_reset_state(SysState *__restrict__ state, int n, double dt0, double alpha0) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    SysState st;
    st.dt = dt0;
    st.alpha = alpha0;
    st.npos = 0;
    st.steps = 0;
    st.status = SMR_SYS_RUNNING;
    st.pad = 0;
    st.fmax = 0.0;
    for (int q = 0; q < 6; q++) st.E[q] = 0.0;
    state[s] = st;
}

// Bits 63-48: the launch slice of position q (slices keep their places);
// 47-32: 65535 - atom count; 31-0: the energy per atom as an order-reversing
// integer (higher energy, smaller key; non-finite energies first).  A stable
// ascending sort then gives each slice larger systems first and, among
// systems of one size, the highest energy per atom first, ties in batch
// order.
__global__ void k_rank_keys(DevTables T, const SysState *__restrict__ state, const int *__restrict__ work_base, int n,
                            RankSlices rs, unsigned long long *__restrict__ keys, int *__restrict__ vals) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int s = work_base[q];
    int slice = 0;
    while (slice + 1 < rs.n && q >= rs.first[slice + 1]) slice++;
    const int natoms = T.topo[T.sys[s].topo].natoms;
    const SysState st = state[s];
    double E = 0.0;
    for (int k = 0; k < 6; k++) E += st.E[k];
    const float e = (float)(E / natoms);
    unsigned ek = 0u;
    if (isfinite(e)) {
        const unsigned u = __float_as_uint(e);
        const unsigned asc = (u & 0x80000000u) ? ~u : (u | 0x80000000u);   // ascending with e
        ek = ~asc;
        if (ek == 0u) ek = 1u;   // 0 is kept for non-finite energies
    }
    keys[q] = ((unsigned long long)slice << 48) | ((unsigned long long)(65535 - natoms) << 32) | ek;
    vals[q] = s;
}

hipError_t launch_reset_state(SysState *state, int n, double dt0, double alpha0, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reset_state, dim3((n + 255) / 256), dim3(256), 0, st, state, n, dt0, alpha0);
    return hipGetLastError();
}

hipError_t rank_work_list(const DevTables &T, const SysState *state, const int *work_base, int n, const RankSlices &rs,
                          RankBuffers &rb, int *work, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipError_t e;
    if (rb.n != n) {
        size_t tmp = 0;
        e = hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, (const unsigned long long *)nullptr,
                                               (unsigned long long *)nullptr, (const int *)nullptr, (int *)nullptr, n, 0,
                                               64, st);
        if (e != hipSuccess) return e;
        if ((e = rb.keys.alloc(2 * sizeof(unsigned long long) * n)) != hipSuccess) return e;
        if ((e = rb.vals.alloc(sizeof(int) * n)) != hipSuccess) return e;
        if ((e = rb.tmp.alloc(tmp)) != hipSuccess) return e;
        rb.tmp_bytes = tmp;
        rb.n = n;
    }
    unsigned long long *kin = (unsigned long long *)rb.keys.ptr, *kout = kin + n;
    hipLaunchKernelGGL(k_rank_keys, dim3((n + 255) / 256), dim3(256), 0, st, T, state, work_base, n, rs, kin,
                       (int *)rb.vals.ptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t tmp = rb.tmp_bytes;
    return hipcub::DeviceRadixSort::SortPairs(rb.tmp.ptr, tmp, kin, kout, (const int *)rb.vals.ptr, work, n, 0, 64, st);
}

}  // namespace smr
