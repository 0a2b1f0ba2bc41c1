// Kernel timing with HIP events (see nslx_prof_* in nslx.h).
#pragma once

#include <string>

#include <hip/hip_ext.h>

#include "common.hpp"

namespace nslx {

// Times one kernel class instance on stream s when timing is on.  The launch
// inside the scope goes through launch_timed(), which hands the scope's
// event pair to hipExtLaunchKernelGGL: the runtime records them at the
// kernel's own start and end, so host gaps between launches (the solvers'
// polls) are not counted.
class ProfScope {
   public:
    ProfScope(const char* name, int prec, int mode, hipStream_t s);
    // a class named in full (kernels other than the hop)
    ProfScope(const std::string& full_name, hipStream_t s);
    ~ProfScope();
    ProfScope(const ProfScope&) = delete;
    ProfScope& operator=(const ProfScope&) = delete;
    // the event pair for the scope's kernel, once (false when timing is off)
    bool take_events(hipEvent_t* start, hipEvent_t* stop);
    // algorithmic flops and bytes of the scope's launch (summed per class
    // over the timed launches, reported beside their time)
    void work(double flops, double bytes) { flops_ = flops; bytes_ = bytes; }

   private:
    bool active_;
    bool used_ = false;
    std::string name_;
    void* a_ = nullptr;
    void* b_ = nullptr;
    double flops_ = 0.0, bytes_ = 0.0;
    void begin();
    ProfScope* outer_ = nullptr;
};

// innermost live ProfScope of this thread (null if none)
ProfScope* prof_current();

bool prof_on();
int prof_every();
void prof_enable(int every);
void prof_reset();
std::string prof_report(int* nclasses);

// Launch a kernel; inside an active ProfScope its events bracket the kernel.
template <class F, class... Args>
void launch_timed(F kernel, dim3 grid, dim3 block, size_t shmem, hipStream_t s, Args... args) {
    hipEvent_t e0, e1;
    ProfScope* p = prof_current();
    if (p && p->take_events(&e0, &e1))
        hipExtLaunchKernelGGL(kernel, grid, block, (uint32_t)shmem, s, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
}

}  // namespace nslx
