// Device side of multi-design batches: the job table and work list of the
// persistent batch kernel.  The kernel itself is generated per batch (each
// design's own item body compiled in: the gate tile body at its variant, the
// word engine's generated code or its interpreter) and compiled at run time,
// so this header, like the two it includes, has no host-only dependency.
#pragma once
#include "word_device.hpp"
#include "gate_device.hpp"

namespace rrtl {
namespace batch {

// One design's kernel arguments; the generated kernel dispatches on the job
// index, so the table carries no kind tag.
struct JobDev {
  union U {
    gate::TileArgs g;
    word::WordArgs w;
  } u;
};

struct BatchArgs {
  const JobDev* jobs;
  const uint2* items;   // (job, item), most expensive first
  uint32_t n_items;
  uint32_t* counter;    // next item; cleared by batch_reset_kernel before every launch
};

typedef const __attribute__((address_space(4))) JobDev const_job;

// The persistent loop: each workgroup takes the next item from the shared
// counter until the list is drained; `run(job, item, args)` runs one item.
// The item index is made wave-uniform, so every wave of the workgroup leaves
// the loop together; the barrier at the end frees LDS for the next item.
// The job table is read through the constant address space (it does not
// change during the launch): with the bodies inlined, their argument reads
// become scalar loads that stores of the simulation cannot clobber, as for a
// kernel's own arguments (through a plain global pointer they are vector
// loads repeated after every store: the 20k-gate tile body ran 45 % slower).
template <class Run>
__device__ __forceinline__ void batch_loop(const BatchArgs& B, Run&& run) {
  __shared__ uint32_t s_item;
  const_job* jobs = (const_job*)B.jobs;  // address-space cast
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(B.counter, 1u);
    __syncthreads();
    const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (it >= B.n_items) break;
    const uint2 wi = B.items[it];
    const uint32_t job = __builtin_amdgcn_readfirstlane(wi.x);
    run(job, __builtin_amdgcn_readfirstlane(wi.y), *(const JobDev::U*)&jobs[job].u);
    __syncthreads();
  }
}

}  // namespace batch
}  // namespace rrtl
