// launch.h -- one kernel body, two launch forms.
//
// A body B provides kThreads, a by-value argument struct Args and
//   static __device__ void run(const Args& a, int blk, int nblk)
// where blk is the workgroup's index among the nblk workgroups of its system.
// k_one launches it for one system (Args by value); k_many launches it for a
// batch of systems in ONE grid: workgroup b runs work item map[b] =
// (system, blk, nblk, 0) with that system's Args read from device memory
// (uniform per workgroup, so they load through the scalar cache).
#pragma once
#include <hip/hip_runtime.h>
#include <string.h>

#include <vector>

namespace rbmd {

template <class B>
__global__ __launch_bounds__(B::kThreads) void k_one(typename B::Args a)
{
    B::run(a, (int)blockIdx.x, (int)gridDim.x);
}

template <class B>
__global__ __launch_bounds__(B::kThreads) void k_many(const typename B::Args* __restrict__ a,
                                                      const int4* __restrict__ map)
{
    const int4 w = map[blockIdx.x];
    B::run(a[w.x], w.y, w.z);
}

// Device-resident per-system arguments and work map of one batched phase.
struct BatchPhase {
    void* args = nullptr;  // Args[nsys]
    int4* map = nullptr;   // work items
    size_t args_bytes = 0;
    int items = 0;
    size_t lds = 0;
    std::vector<char> host_args;  // last uploaded copy
    std::vector<int> host_blocks;
    void release()
    {
        if (args) (void)hipFree(args);
        if (map) (void)hipFree(map);
        args = nullptr;
        map = nullptr;
        args_bytes = 0;
        items = 0;
        host_args.clear();
        host_blocks.clear();
    }
};

// Upload the arguments and (when the block counts changed) the work map of a
// batched phase; synchronous, done outside graph capture.  Work items are
// ordered system by system.
template <class Args>
hipError_t phase_upload(BatchPhase& ph, const std::vector<Args>& args,
                        const std::vector<int>& blocks, size_t lds)
{
    const size_t bytes = args.size() * sizeof(Args);
    hipError_t e = hipSuccess;
    if (bytes > ph.args_bytes) {
        if (ph.args) (void)hipFree(ph.args);
        ph.args = nullptr;
        ph.args_bytes = 0;
        ph.host_args.clear();
        if ((e = hipMalloc(&ph.args, bytes)) != hipSuccess) return e;
        ph.args_bytes = bytes;
    }
    const char* src = reinterpret_cast<const char*>(args.data());
    if (ph.host_args.size() != bytes || memcmp(ph.host_args.data(), src, bytes) != 0) {
        if ((e = hipMemcpy(ph.args, src, bytes, hipMemcpyHostToDevice)) != hipSuccess) return e;
        ph.host_args.assign(src, src + bytes);
    }
    if (blocks != ph.host_blocks) {
        std::vector<int4> m;
        for (size_t r = 0; r < blocks.size(); ++r)
            for (int b = 0; b < blocks[r]; ++b) m.push_back(make_int4((int)r, b, blocks[r], 0));
        if (ph.map) (void)hipFree(ph.map);
        ph.map = nullptr;
        ph.items = 0;
        if (!m.empty()) {
            if ((e = hipMalloc((void**)&ph.map, m.size() * sizeof(int4))) != hipSuccess)
                return e;
            if ((e = hipMemcpy(ph.map, m.data(), m.size() * sizeof(int4),
                               hipMemcpyHostToDevice)) != hipSuccess)
                return e;
        }
        ph.items = (int)m.size();
        ph.host_blocks = blocks;
    }
    ph.lds = lds;
    return hipSuccess;
}

template <class B>
void phase_launch(const BatchPhase& ph, hipStream_t s)
{
    if (ph.items > 0)
        k_many<B><<<ph.items, B::kThreads, ph.lds, s>>>((const typename B::Args*)ph.args, ph.map);
}

}  // namespace rbmd
