"""Ensemble workflows on a store with NSL's node layout (nsl_amd.store):
the HMC Markov chain written state by state and resumed from the store, and
the two-point correlator measured over the stored configurations.

``generate_chain`` is HMC::generate<AllStates>(state, Nconf, saveFrequency,
baseNode) with its file side (src/NSL/MarkovChain/HMC.tpp:57-155); the
trajectories run on the device (nslx_hmc_generate).  ``measure_two_point``
is TwoPointCorrelator::measure() (src/NSL/Measurements/Hubbard/
twoPointCorrelationFunction.tpp:241-315): the non-interacting correlator,
then every stored configuration, with the reference's skip / overwrite
rules; configurations are measured ``batch`` at a time as the fields of one
fermion matrix (nslx_correlator_two_point_fields), which solves all their
point sources in one time-slice-preconditioned CG.
"""

import ctypes

from . import store as _store


def _torch():
    import torch
    return torch


def generate_chain(action, store, phi0, nconf, traj_length, nmd, seed, save_frequency=1, base_node="markovChain",
                   block=64):
    """Generate the chain ``base_node/0 .. base_node/nconf-1`` in store.

    As HMC.tpp:74-98: if base_node already holds states, the chain resumes
    from the largest id; if that record cannot be read it is deleted and the
    chain resumes from the one before (:86-95); otherwise phi0 (device or
    host, (Nt, Nx)) is written as state 0 (:96-98).  States n = nstart+1 ..
    nconf-1 are generated (saveFrequency trajectories each) and written as
    they come, ``block`` states per device call.  A resumed chain continues
    bit for bit: the trajectory that makes markov time m+1 draws its momenta
    and uniform from Philox counters 2m and 2m+1 of seed (nslx_hmc_generate),
    and the markov time is read back from the record (the reference uses the
    node id, which equals it for saveFrequency 1).

    Returns the number of states written by this call."""
    torch = _torch()
    nstart = 0
    if store.exist(base_node) and store.list(base_node):
        _, nstart = store.min_max_configs(base_node)
    if store.exist(f"{base_node}/{nstart}"):
        try:
            _, phi_h, st = _store.read_markov_state(store, base_node, nstart)
        except (OSError, KeyError, ValueError):
            store.delete(f"{base_node}/{nstart}")
            nstart -= 1
            _, phi_h, st = _store.read_markov_state(store, base_node, nstart)
        phi = torch.as_tensor(phi_h).to(device="cuda", dtype=torch.complex128).contiguous()
        markov_time = st["markov_time"]
        written = 0
    else:
        phi = torch.as_tensor(phi0).to(device="cuda", dtype=torch.complex128).contiguous().clone()
        states, _ = action.generate(phi, 1, traj_length, nmd, seed, save_frequency, 0, all_states=True,
                                    keep_configs=False)
        _store.write_markov_state(store, f"{base_node}/0", phi, states[0])
        markov_time = states[0]["markov_time"]
        written = 1
    n = nstart + 1
    while n < nconf:
        k = min(int(block), nconf - n)
        # k + 1 states: [0] is the start (already stored), [1..k] are new
        states, cfgs = action.generate(phi, k + 1, traj_length, nmd, seed, save_frequency, markov_time,
                                       all_states=True, keep_configs=True)
        for i in range(1, k + 1):
            _store.write_markov_state(store, f"{base_node}/{n + i - 1}", cfgs[i], states[i])
        written += k
        n += k
        markov_time = states[k]["markov_time"]
        phi = cfgs[k].clone()
    return written


def _skip(store, overwrite, node):
    # twoPointCorrelationFunction.tpp:98-110 (skip_): recompute when asked
    # to overwrite or when the node does not exist
    return (not overwrite) and store.exist(node)


def measure_two_point(store, lat, Nt, beta, mu=0.0, species=None, number_time_sources=None, basenode="",
                      overwrite=False, tol=1e-12, precision=None, batch=16, maxIter=2000, ctx=None):
    """TwoPointCorrelator<..., HubbardExp>::measure() over the store.

    Writes ``<basenode>/NonInteracting/correlators/single/<species>`` (phi =
    0, one time source, :248-271) and, for every id of
    ``<basenode>/markovChain``, ``<basenode>/markovChain/<id>/correlators/
    single/<species>`` from ``<basenode>/markovChain/<id>/phi`` with
    number_time_sources time sources (default Nt; :273-313), each a
    (Nt, Nx, Nx) complex128 dataset corr[t, y, x].  Existing correlators are
    kept unless overwrite.  The solves go to |b - MM^dag x| <= tol |b| per
    point source (double-double), preconditioner in ``precision`` (F64 by
    default).  Returns the list of ids measured by this call."""
    torch = _torch()
    from . import F64, HubbardExp, Particle, SolveReport, _lib, _ptr, _stream
    species = Particle if species is None else int(species)
    precision = F64 if precision is None else int(precision)
    Nx = lat.sites()
    name = "particle" if species == Particle else "hole"
    nts = int(number_time_sources or Nt)

    def measure(phis, ntsrc):
        fm = HubbardExp(lat, Nt, beta, mu, species, ctx)
        fm.populate(phis)
        nf = phis.shape[0]
        corr = torch.empty((nf, Nt, Nx, Nx), dtype=torch.complex128, device="cuda")
        rep = SolveReport()
        st = _lib.lib().nslx_correlator_two_point_fields(fm._h, int(ntsrc), float(tol), int(maxIter), precision,
                                                         0, _ptr(corr), ctypes.byref(rep), _stream())
        _lib.check(st, allow=(_lib.OK,))
        return corr

    node = f"{basenode}/NonInteracting/correlators/single/{name}"
    if not _skip(store, overwrite, node):
        zero = torch.zeros((1, Nt, Nx), dtype=torch.complex128, device="cuda")
        store.write(node, measure(zero, 1)[0])

    chain = f"{basenode}/markovChain"
    lo, hi = store.min_max_configs(chain)
    todo = []
    trim = False
    for cfg in range(lo, hi + 1):
        cnode = f"{chain}/{cfg}/correlators/single/{name}"
        if _skip(store, overwrite, cnode):
            trim = True
            continue
        if trim:
            # :296-299 deletes the node of the next configuration to compute
            # (absent unless overwrite, in which case nothing was skipped)
            if store.exist(cnode):
                store.delete(cnode)
            trim = False
        if store.exist(f"{chain}/{cfg}/phi"):
            todo.append(cfg)
    bsz = max(1, int(batch))
    for i in range(0, len(todo), bsz):
        ids = todo[i:i + bsz]
        phis = torch.stack([torch.as_tensor(store.read(f"{chain}/{c}/phi")).to(torch.complex128) for c in ids])
        corr = measure(phis.to("cuda").contiguous(), nts)
        for j, c in enumerate(ids):
            store.write(f"{chain}/{c}/correlators/single/{name}", corr[j])
    return todo
