"""The ensemble workflows on the device: an HMC chain written to a store and
resumed from it (HMC.tpp:57-155), and the two-point correlator measured over
the stored configurations (twoPointCorrelationFunction.tpp:241-315), against
the oracle.

Tolerances: configurations and actions of the stored chain against the
oracle's chain to 1e-9 (as the in-memory chain test); a resumed chain equal
bit for bit to the uninterrupted one; each configuration's correlator
against the oracle's measure(N) (its CG at eps 1e-12) to 1e-9 relative."""

import math

import numpy as np
import pytest
import torch

import nsl_amd as nsl
from nsl_amd import ensemble, store as st
from oracle import correlator as ocorr, fermion as F, hmc as ohmc, lattice as olat, rng

pytestmark = pytest.mark.gpu

L, NT, BETA, U = 3, 16, 4.0, 2.0
NX = 2 * L * L
TRAJ, NMD, SEED = 1.0, 10, 3000


def phi0():
    return torch.as_tensor(rng.real_field(1000, 0, (NT, NX), math.sqrt(BETA / NT * U)))


def action():
    return nsl.HubbardAction(nsl.Lattice.honeycomb(L, L), NT, BETA, U)


def chain_in(path, nconf, block=64):
    s = st.DirStore(path)
    ensemble.generate_chain(action(), s, phi0(), nconf, TRAJ, NMD, SEED, block=block)
    return s


def same_records(a, b, ids):
    for i in ids:
        for d in ("phi", "actVal", "acceptanceProbability", "acceptanceRate", "markovTime"):
            assert np.array_equal(a.read(f"markovChain/{i}/{d}"), b.read(f"markovChain/{i}/{d}")), (i, d)


def test_stored_chain_matches_oracle(tmp_path):
    s = chain_in(tmp_path / "a", 6, block=2)
    assert s.min_max_configs("markovChain") == (0, 5)
    Ao = ohmc.HubbardAction(olat.honeycomb(L, L), NT, BETA, U)
    ref = ohmc.hmc_generate(Ao, phi0(), 6, TRAJ, NMD, SEED)
    for i, o in enumerate(ref):
        _, p, stt = st.read_markov_state(s, "markovChain", i)
        assert np.abs(p - o["phi"].numpy()).max() < 1e-9 * np.abs(o["phi"].numpy()).max(), i
        assert abs(stt["S_re"] - o["S"].real) < 1e-9 * abs(o["S"].real), i
        assert stt["markov_time"] == o["markov_time"] and bool(stt["accepted"]) == bool(o["accepted"]), i


def test_chain_resumes_bit_for_bit(tmp_path):
    """3 states, then the same call asked for 7: the store continues from id
    2 and equals a 7-state chain written in one go, record for record."""
    whole = chain_in(tmp_path / "whole", 7)
    part = chain_in(tmp_path / "part", 3)
    assert ensemble.generate_chain(action(), part, phi0(), 7, TRAJ, NMD, SEED) == 4
    same_records(whole, part, range(7))
    # nothing left to do: no state written
    assert ensemble.generate_chain(action(), part, phi0(), 7, TRAJ, NMD, SEED) == 0


def test_chain_resume_falls_back_over_a_damaged_record(tmp_path):
    """The last record unreadable (HMC.tpp:86-95): it is deleted, the chain
    resumes from the state before it and rewrites it identically."""
    whole = chain_in(tmp_path / "whole", 6)
    part = chain_in(tmp_path / "part", 4)
    part.delete("markovChain/3/actVal")
    assert ensemble.generate_chain(action(), part, phi0(), 6, TRAJ, NMD, SEED) == 3
    same_records(whole, part, range(6))


def oracle_corr(phi, ntsrc):
    o = F.HubbardExp(olat.honeycomb(L, L), NT, BETA)
    o.populate(torch.as_tensor(phi))
    corr = ocorr.two_point(o, NT, NX, ntsrc)
    return corr[0] if isinstance(corr, tuple) else corr


def test_measure_over_the_store_matches_oracle(tmp_path):
    s = chain_in(tmp_path / "m", 5)
    ids = ensemble.measure_two_point(s, nsl.Lattice.honeycomb(L, L), NT, BETA, number_time_sources=4, batch=3)
    assert ids == list(range(5))
    free = s.read("/NonInteracting/correlators/single/particle")
    ref = oracle_corr(np.zeros((NT, NX), dtype=np.complex128), 1).numpy()
    assert free.shape == (NT, NX, NX)
    assert np.abs(free - ref).max() < 1e-9 * np.abs(ref).max()
    for i in ids:
        got = s.read(f"markovChain/{i}/correlators/single/particle")
        ref = oracle_corr(s.read(f"markovChain/{i}/phi"), 4).numpy()
        assert np.abs(got - ref).max() < 1e-9 * np.abs(ref).max(), i


def test_measure_skip_overwrite_and_batching(tmp_path):
    """Existing correlators are kept unless overwrite (skip_,
    twoPointCorrelationFunction.tpp:98-110); the batch size does not change
    the results beyond rounding; hole correlators go to their own nodes."""
    s = chain_in(tmp_path / "s", 4)
    lat = nsl.Lattice.honeycomb(L, L)
    ensemble.measure_two_point(s, lat, NT, BETA, number_time_sources=2, batch=4)
    first = {i: s.read(f"markovChain/{i}/correlators/single/particle") for i in range(4)}
    s.write("markovChain/2/correlators/single/particle", np.zeros((NT, NX, NX), dtype=np.complex128))
    assert ensemble.measure_two_point(s, lat, NT, BETA, number_time_sources=2) == []
    assert not s.read("markovChain/2/correlators/single/particle").any()
    assert ensemble.measure_two_point(s, lat, NT, BETA, number_time_sources=2, overwrite=True, batch=1) == [0, 1, 2, 3]
    for i in range(4):
        again = s.read(f"markovChain/{i}/correlators/single/particle")
        assert np.abs(again - first[i]).max() < 1e-12 * np.abs(first[i]).max(), i
    # a configuration without a correlator is measured, the others kept
    s.delete("markovChain/1/correlators")
    assert ensemble.measure_two_point(s, lat, NT, BETA, number_time_sources=2) == [1]
    ensemble.measure_two_point(s, lat, NT, BETA, species=nsl.Hole, number_time_sources=2)
    assert s.exist("markovChain/3/correlators/single/hole") and s.exist("/NonInteracting/correlators/single/hole")
