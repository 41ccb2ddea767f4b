// group_members.hip -- particles grouped by structure on the device: the relabelling of
// select_with_MS_method_1319 (src/compute_neiKDtree_mod.py:376-379, whereIam_parts ->
// node_to_struct) and the member grouping of compute_adaptahop_131 (:109-119, argsort of the
// labels and np.unique's first indices and counts), as one radix sort of (label, particle)
// pairs and a boundary pass that writes CSR offsets for every label.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include <rocprim/device/device_radix_sort.hpp>

#include "ahop_common.h"

namespace ahop {

namespace {

// labels[i] = node_to_struct[node[i] - 1] for node[i] > 0 (0 stays 0), or node[i] itself
// without a map; ids outside their range are counted in bad
__global__ void k_relabel(const int32_t *__restrict__ node, int64_t n,
                          const int32_t *__restrict__ node_to_struct, int32_t nnodes,
                          int32_t nlab, uint32_t *__restrict__ labels, int32_t *__restrict__ iota,
                          int *__restrict__ bad)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = node[i];
    int32_t l = 0;
    if (v < 0 || v > nnodes) {
        atomicAdd(bad, 1);
    } else if (v > 0) {
        l = node_to_struct ? node_to_struct[v - 1] : v;
        if (l < 1 || l > nlab) {
            atomicAdd(bad, 1);
            l = 0;
        }
    }
    labels[i] = (uint32_t)l;
    iota[i] = (int32_t)i;
}

// offsets[l] = first sorted position with label >= l, for l = 0 .. nlab + 1 (every label is
// written once: by the position where the sorted labels first pass it, or by the end)
__global__ void k_label_offsets(const uint32_t *__restrict__ sorted, int64_t n, int32_t nlab,
                                int64_t *__restrict__ offsets)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i > n) return;
    const int64_t lab = i < n ? (int64_t)sorted[i] : (int64_t)nlab + 1;
    const int64_t prev = i > 0 ? (int64_t)sorted[i - 1] : -1;
    for (int64_t l = prev + 1; l <= lab; ++l) offsets[l] = i;
}

__global__ void k_copy_labels(const uint32_t *__restrict__ labels, int64_t n,
                              int32_t *__restrict__ out)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int32_t)labels[i];
}

int bits_for(int64_t v)
{
    int b = 1;
    while (b < 32 && (v >> b) != 0) ++b;
    return b;
}

}  // namespace

float group_members(const int32_t *d_node, int64_t n, const int32_t *d_node_to_struct,
                    int32_t nnodes, int32_t nlab, int32_t *d_labels_out, int32_t *d_pids,
                    int64_t *d_offsets, hipStream_t s)
{
    if (n < 1 || n > INT32_MAX) throw std::runtime_error("need 1 <= npart <= 2^31-1");
    if (nnodes < 0 || nlab < 0) throw std::runtime_error("nnodes and nstruct must be >= 0");
    if (!d_node_to_struct && nlab != nnodes)
        throw std::runtime_error("without a node-to-structure map the labels are the node ids");
    static DevBuf keys, keys_sorted, iota, temp, badbuf;
    uint32_t *k0 = (uint32_t *)keys.ensure(sizeof(uint32_t) * n);
    uint32_t *k1 = (uint32_t *)keys_sorted.ensure(sizeof(uint32_t) * n);
    int32_t *io = (int32_t *)iota.ensure(sizeof(int32_t) * n);
    int *d_bad = (int *)badbuf.ensure(sizeof(int));
    hipEvent_t e0, e1;
    AHOP_HIP(hipEventCreate(&e0));
    AHOP_HIP(hipEventCreate(&e1));
    const unsigned nb = (unsigned)((n + 255) / 256), nb1 = (unsigned)((n + 256) / 256);
    AHOP_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), s));
    AHOP_HIP(hipEventRecord(e0, s));
    hipLaunchKernelGGL(k_relabel, dim3(nb), dim3(256), 0, s, d_node, n, d_node_to_struct, nnodes,
                       nlab, k0, io, d_bad);
    size_t tb = 0;
    const int end_bit = bits_for(nlab);
    AHOP_HIP(rocprim::radix_sort_pairs(nullptr, tb, k0, k1, io, d_pids, (size_t)n, 0, end_bit, s));
    AHOP_HIP(rocprim::radix_sort_pairs(temp.ensure(tb), tb, k0, k1, io, d_pids, (size_t)n, 0,
                                       end_bit, s));
    hipLaunchKernelGGL(k_label_offsets, dim3(nb1), dim3(256), 0, s, k1, n, nlab, d_offsets);
    if (d_labels_out)
        hipLaunchKernelGGL(k_copy_labels, dim3(nb), dim3(256), 0, s, k0, n, d_labels_out);
    AHOP_HIP(hipEventRecord(e1, s));
    int bad = 0;
    AHOP_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
    AHOP_HIP(hipStreamSynchronize(s));
    float ms = 0;
    AHOP_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (bad)
        throw std::runtime_error(std::to_string(bad) +
                                 " node ids outside [0, nnodes] or structure ids outside [1, nstruct]");
    return ms;
}

}  // namespace ahop
