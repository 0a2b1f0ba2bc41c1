// Kernel timing with HIP events (see nslx_prof_* in nslx.h).
#pragma once

#include <string>

#include "common.hpp"

namespace nslx {

// Records an event pair around the scope on stream s when timing is on.
class ProfScope {
   public:
    ProfScope(const char* name, int prec, int mode, hipStream_t s);
    ~ProfScope();
    ProfScope(const ProfScope&) = delete;
    ProfScope& operator=(const ProfScope&) = delete;

   private:
    hipStream_t s_;
    bool active_;
    std::string name_;
    void* a_ = nullptr;
    void* b_ = nullptr;
};

bool prof_on();
void prof_enable(bool on);
void prof_reset();
std::string prof_report(int* nclasses);

}  // namespace nslx
