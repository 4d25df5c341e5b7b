/*
 * sa_plane_sort.hip -- the HDP emission plane's column groups, built on the device.
 *
 * The column pass of the emission plane stages one density row in LDS per workgroup,
 * so it takes the band's columns grouped by (stream lane, density row).  Each column
 * (task, x) gets the key lane * (nrow + 1) + row + 1 (row -1: column 0, or a k-mer
 * with no density row); a stable radix sort by key keeps, inside a key, the lane's
 * task order and ascending x, and a run-length pass gives each key's column count,
 * from which the host cuts groups of at most kColumnGroup columns.  Only the runs
 * (at most 4 * (nrow + 1) keys and counts) come back to the host.
 */
#include "sa_kernels.h"

#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_run_length_encode.hpp>

namespace sa {

/* one workgroup per plane task: its columns' keys and (task, x) values */
__global__ void __launch_bounds__(256) plane_keys_kernel(const PlaneTask *__restrict__ pt,
                                                         const int32_t *__restrict__ kid,
                                                         const int32_t *__restrict__ hrow, int32_t nrow,
                                                         uint32_t *__restrict__ keys, uint64_t *__restrict__ vals) {
    const PlaneTask t = pt[blockIdx.x];
    const uint32_t base = (uint32_t)t.lane * (uint32_t)(nrow + 1);
    for (int32_t x = threadIdx.x; x <= t.lX; x += 256) {
        int32_t r = -1;
        if (x > 0) {
            r = hrow[kid[t.pos_off + x - 1]];
            if (r < 0 || r >= nrow) r = -1;
        }
        keys[t.col0 + x] = base + (uint32_t)(r + 1);
        vals[t.col0 + x] = ((uint64_t)(uint32_t)x << 32) | (uint32_t)t.task; /* as int32 pairs: task, x */
    }
}

static int key_bits(int32_t nrow) {
    const uint64_t top = (uint64_t)kPlaneLanes * (uint64_t)(nrow + 1);
    int b = 1;
    while (b < 32 && ((uint64_t)1 << b) < top) b++;
    return b;
}

size_t plane_sort_temp_bytes(int64_t n_cols, int32_t nrow) {
    size_t a = 0, b = 0;
    const int n = (int)n_cols;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, a, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                           (const uint64_t *)nullptr, (uint64_t *)nullptr, n, 0,
                                           key_bits(nrow)) != hipSuccess)
        return 0;
    if (hipcub::DeviceRunLengthEncode::Encode(nullptr, b, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                              (int32_t *)nullptr, (int32_t *)nullptr, n) != hipSuccess)
        return 0;
    return std::max(a, b);
}

hipError_t launch_plane_sort(const PlaneTask *pt, int n_tasks, int64_t n_cols, const int32_t *kid,
                             const int32_t *hrow, int32_t nrow, uint32_t *keys, uint32_t *keys_sorted,
                             uint64_t *vals, uint64_t *entries, uint32_t *run_keys, int32_t *run_counts,
                             int32_t *n_runs, void *temp, size_t temp_bytes, hipStream_t st) {
    if (n_tasks <= 0 || n_cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(plane_keys_kernel, dim3(n_tasks), dim3(256), 0, st, pt, kid, hrow, nrow, keys, vals);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n = (int)n_cols;
    size_t tb = temp_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(temp, tb, keys, keys_sorted, vals, entries, n, 0, key_bits(nrow), st);
    if (e != hipSuccess) return e;
    tb = temp_bytes;
    return hipcub::DeviceRunLengthEncode::Encode(temp, tb, keys_sorted, run_keys, run_counts, n_runs, n, st);
}

} // namespace sa
