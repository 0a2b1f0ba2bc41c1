// HMC pieces on the device (hmc.hip) and the host RNG / lattice builders
// (host.cpp) shared by the C ABI.
#pragma once

#include "engine.hpp"

namespace nslx {

void hmc_action(FermionMatrix& fm, const void* phi, double U, int nchunks, double* out2, hipStream_t s);
void hmc_force(FermionMatrix& fm, const void* phi, double U, void* out, hipStream_t s);
void hmc_trajectory(FermionMatrix& fm, void* phi, const void* mom, double* S0, double U, double T, int nmd, double uni,
                    int nchunks, nslx_hmc_report* rep, hipStream_t s);
void hmc_chain(FermionMatrix& fm, void* phi, double U, double T, int nmd, int64_t ntraj, uint64_t seed, int nchunks,
               nslx_hmc_report* reps, hipStream_t s);
// HMC::generate<AllStates | LastState>(state, Nconf, saveFrequency) without
// the file I/O: states (host, nconf or 1) and configs (device, nconf fields)
// may be null.
void hmc_generate(FermionMatrix& fm, void* phi, int64_t mt0, double U, double T, int nmd, int64_t nconf,
                  int64_t save, int chain, uint64_t seed, int nchunks, void* configs, nslx_markov_state* states,
                  hipStream_t s);

// Philox-4x32-10 (Salmon et al. SC'11); stream layout documented in nslx.h.
void host_philox(uint64_t seed, uint64_t offset, int64_t nblocks, uint32_t* out4);
void host_uniform(uint64_t seed, uint64_t offset, int64_t n, double* out);
void host_normal(uint64_t seed, uint64_t offset, int64_t n, double* out);

// Lattices (host); K row-major nx*nx.
void lattice_honeycomb(int L0, int L1, double kappa, double* K, double* sites);
void lattice_ring(int n, double kappa, double* K, double* sites);
void lattice_nanotube(int n, int cells, double kappa, double* K, double* sites);

}  // namespace nslx
