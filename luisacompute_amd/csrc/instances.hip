// instances.hip -- the DSL's in-kernel instance ops over an accel's device-side instance table
// (RAY_TRACING_INSTANCE_TRANSFORM / _USER_ID / _VISIBILITY_MASK and their SET_ forms,
// include/luisa/ast/op.h:290-297; CPU backend semantics src/rust/luisa_compute_backend_impl/src/
// cpu/accel.rs:537-579 and cpu/stream.rs:582-660), run as batches: item i reads or writes field
// `field` of instance idx[i]. Reads see the latest writes; writes mark the instance dirty and
// reach traversal at the accel's next build, where the host folds them into its instance set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "lc_hip_internal.h"

namespace lchip {

namespace {

__global__ void k_instance_read(const InstanceState *st, uint32_t count, uint32_t field, const uint32_t *idx,
                                uint32_t n, void *out, uint32_t *err) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t k = idx[i];
        const bool ok = k < count;
        if (!ok) atomicOr(err, 1u);
        if (field == kInstanceTransform) {
            float *o = static_cast<float *>(out) + 12 * (uint64_t)i;
            for (int c = 0; c < 12; c++) o[c] = ok ? st[k].affine[c] : 0.0f;
        } else {
            uint32_t *o = static_cast<uint32_t *>(out);
            o[i] = ok ? (field == kInstanceUserId ? st[k].user_id : st[k].vis_mask) : 0u;
        }
    }
}

__global__ void k_instance_write(InstanceState *st, uint32_t count, uint32_t field, const uint32_t *idx, uint32_t n,
                                 const void *in, uint32_t *err) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t k = idx[i];
        if (k >= count) {
            atomicOr(err, 1u);
            continue;
        }
        if (field == kInstanceTransform) {
            const float *v = static_cast<const float *>(in) + 12 * (uint64_t)i;
            for (int c = 0; c < 12; c++) st[k].affine[c] = v[c];
        } else if (field == kInstanceUserId) {
            st[k].user_id = static_cast<const uint32_t *>(in)[i];
        } else {
            st[k].vis_mask = static_cast<const uint32_t *>(in)[i];
        }
        atomicOr(&st[k].dirty, 1u << field);
    }
}

} // namespace

hipError_t instance_ops(hipStream_t stream, InstanceState *st, uint32_t count, uint32_t field, bool write,
                        const uint32_t *idx, uint32_t n, void *data, uint32_t *err) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = std::min<uint32_t>((n + 255) / 256, 4096);
    if (write)
        hipLaunchKernelGGL(k_instance_write, dim3(grid), dim3(256), 0, stream, st, count, field, idx, n,
                           (const void *)data, err);
    else
        hipLaunchKernelGGL(k_instance_read, dim3(grid), dim3(256), 0, stream, (const InstanceState *)st, count, field,
                           idx, n, data, err);
    return hipGetLastError();
}

} // namespace lchip
