// Batched persistent projection kernels (dh_lra_batched): one per tile class
// and operand type (BF16 or FP16).
#include "batch.h"
#include "rbf16_kernels.h"

namespace {

// Persistent workgroups: each takes the next item from *counter until the
// queue is empty; every wave of the workgroup sees the same item and leaves
// the loop together.  An item is 256 rows (128 for d > 128) of one job.
template <int DK, int NCB, int WAVES, int RB, bool F16>
__global__ __launch_bounds__(WAVES * 64, 1) void rbf_project_bf16_persistent(
    const DhProjJob* __restrict__ jobs, const DhProjItem* __restrict__ items, int nitems,
    int* __restrict__ counter) {
    __shared__ int s_item;
    __shared__ ProjLds<DK, NCB, WAVES, RB, false, false> lds;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const DhProjItem w = items[it];
        const DhProjJob& j = jobs[w.job];
        project_rows<DK, NCB, WAVES, F16, RB>(lds, w.rblk, w.cg, j.m, j.nchunks, j.l, j.ncb_pack,
                                                (const bf16x8*)j.xb, (const bf16x8*)j.xb, j.a, j.a,
                                                nullptr, nullptr, j.c2, (const bf16x8*)j.Op, j.y, j.Y,
                                                j.ldy);
    }
}

template <int DK, int NCB, int W, int RB>
int launch_persistent(dh_handle_t h, bool f16, const DhProjJob* jobs, const DhProjItem* items,
                      int nitems, int* counter) {
    auto kern = f16 ? rbf_project_bf16_persistent<DK, NCB, W, RB, true>
                    : rbf_project_bf16_persistent<DK, NCB, W, RB, false>;
    int per_cu = 1;
    DH_CHECK_STATUS(dh_occupancy(h, (const void*)kern, W * 64, &per_cu));
    int grid = max(1, min(nitems, per_cu * h->num_cu));
    kern<<<grid, W * 64, 0, h->stream>>>(jobs, items, nitems, counter);
    DH_CHECK_LAUNCH();
    return 0;
}

template <int DK>
int persistent_ncb(dh_handle_t h, int ncb, bool wide, bool f16, const DhProjJob* jobs,
                   const DhProjItem* items, int nitems, int* counter) {
#define DH_PERS(N)                                                                             \
    case N:                                                                                    \
        if constexpr (DK >= 8)                                                                 \
            return launch_persistent<DK, N, 4, 2>(h, f16, jobs, items, nitems, counter);       \
        else                                                                                   \
            return wide ? launch_persistent<DK, N, 4, 4>(h, f16, jobs, items, nitems, counter) \
                        : launch_persistent<DK, N, 8, 2>(h, f16, jobs, items, nitems, counter);
    switch (ncb) {
        DH_PERS(1)
        DH_PERS(2)
        DH_PERS(3)
        DH_PERS(5)
        DH_PERS(9)
        DH_PERS(17)
        default:
            return dh_fail(DH_ERR_UNSUPPORTED, "rbf_project_bf16: no kernel for %d column blocks",
                           ncb);
    }
#undef DH_PERS
}

}  // namespace

int dh_bf16_project_persistent(dh_handle_t h, DhBf16Class c, const DhProjJob* jobs,
                               const DhProjItem* items, int nitems, int* counter) {
    if (nitems == 0) return 0;
    const bool wide = c.wide != 0, f16 = c.f16 != 0;
    switch (c.dk) {
        case 1: return persistent_ncb<1>(h, c.ncb, wide, f16, jobs, items, nitems, counter);
        case 2: return persistent_ncb<2>(h, c.ncb, wide, f16, jobs, items, nitems, counter);
        case 4: return persistent_ncb<4>(h, c.ncb, wide, f16, jobs, items, nitems, counter);
        case 8: return persistent_ncb<8>(h, c.ncb, false, f16, jobs, items, nitems, counter);
        default: return dh_fail(DH_ERR_UNSUPPORTED, "rbf_project_bf16: feature blocks %d", c.dk);
    }
}
