// Application of the fermion matrix that allocates no device memory.
//
// nslx_fm_apply composes MMdagger / MdaggerM from two hops through a work
// vector that the fermion matrix grows on demand (hipMalloc).  Callers that
// cannot allocate -- code under HIP graph capture, tight serving loops,
// harnesses that forbid it -- use nslx_fm_apply_noalloc instead: its
// intermediate lives in the file-scope device array `scratch` below, one copy
// per GPU, held until the process ends.  Every call on a GPU shares that
// GPU's array, so calls that overlap on one GPU (two streams, two threads)
// race; the array keeps the last call's contents, and nothing here needs it
// to start at zero (the first hop overwrites every element it reads back).
#include <array>
#include <mutex>

#include "engine.hpp"
#include "prof.hpp"

// Intermediate of the two-hop operators: up to NSLX_NOALLOC_SCRATCH_BYTES
// (64 MiB: 227 right-hand sides of 64 x 288 c128 sites).
static __device__ __align__(256) unsigned char scratch[NSLX_NOALLOC_SCRATCH_BYTES];

namespace nslx {
namespace {
std::mutex g_mu;
std::array<void*, 64> g_dev_ptr{};   // device address of `scratch`, per device ordinal
}  // namespace

// Called with `device` current (nslx_ctx_create): looks up this device's
// copy of `scratch` once.  In host code the array's name is only a handle
// for the runtime; kernels and copies get the looked-up address.
void noalloc_init(int device) {
    NSLX_REQUIRE(device >= 0 && device < (int)g_dev_ptr.size(), "noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> g(g_mu);
    if (g_dev_ptr[device]) return;
    void* p = nullptr;
    NSLX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NSLX_REQUIRE(p != nullptr, "noalloc: no device address for the scratch array");
    g_dev_ptr[device] = p;
}

void apply_noalloc(FermionMatrix& fm, int op, const void* in, void* out, int64_t nb, int prec, hipStream_t s) {
    NSLX_REQUIRE(prec == NSLX_F64 || prec == NSLX_F32 || prec == NSLX_BF16X3, "apply_noalloc: bad precision");
    NSLX_REQUIRE(op >= NSLX_OP_M && op <= NSLX_OP_MDAGGERM, "apply_noalloc: bad op");
    NSLX_REQUIRE(in && out, "apply_noalloc: null buffer");
    NSLX_REQUIRE(in != out, "apply_noalloc: in and out must not overlap");
    fm.require_populated();
    NSLX_REQUIRE(nb >= 1 && nb % fm.nfield == 0, "apply_noalloc: nbatch must be a positive multiple of nfield");
    // kernel timing creates HIP events on its first use
    if (prof_on()) throw Error(NSLX_E_STATE, "apply_noalloc: kernel timing (nslx_prof_enable) allocates; turn it off");
    // the hop matrix is built by populate(); building it here would allocate
    if (!fm.has_hop_matrix()) throw Error(NSLX_E_STATE, "apply_noalloc: the fermion matrix has no hop matrix yet");
    if (op == NSLX_OP_M || op == NSLX_OP_MDAGGER) {
        fm.hop(prec, op == NSLX_OP_M ? HOP_M : HOP_MD, in, out, nb, s);
        return;
    }
    const size_t esz = prec == NSLX_F64 ? 16 : 8;
    const size_t need = (size_t)nb * (size_t)fm.len() * esz;
    if (need > (size_t)NSLX_NOALLOC_SCRATCH_BYTES)
        throw Error(NSLX_E_ARG, "apply_noalloc: nbatch * nt * nx * " + std::to_string(esz) + " = " +
                                    std::to_string(need) + " bytes exceeds the " +
                                    std::to_string((size_t)NSLX_NOALLOC_SCRATCH_BYTES) + "-byte scratch array");
    int dev = -1;
    NSLX_HIP(hipGetDevice(&dev));
    void* w = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (dev >= 0 && dev < (int)g_dev_ptr.size()) w = g_dev_ptr[dev];
    }
    if (!w) throw Error(NSLX_E_STATE, "apply_noalloc: scratch array not set up on this device");
    const int first = op == NSLX_OP_MMDAGGER ? HOP_MD : HOP_M;
    const int second = op == NSLX_OP_MMDAGGER ? HOP_M : HOP_MD;
    fm.hop(prec, first, in, w, nb, s);
    fm.hop(prec, second, w, out, nb, s);
}

}  // namespace nslx
