"""Hubbard actions, leapfrog (real force) and one HMC trajectory (CPU oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* U~ = delta*U: src/NSL/Action/Implementations/hubbard.tpp:10-19.
* Gauge action S_G = sum(phi*phi)/(2U~), force -phi/U~:
  hubbardGaugeAction.tpp:59-67.
* Fermion action S_F = -(logDetM_p + logDetM_h), grad = gradLogDetM_p -
  gradLogDetM_h, force = -grad: hubbardFermiAction.tpp:68-113.
* LeapfrogRealForce: src/NSL/Integrator/Impl/leapfrogRealForce.tpp:35-74
  (Im of each force zeroed; half step, Nmd-1 full steps, half step).
* HMC::generate_: src/NSL/MarkovChain/HMC.tpp:185-241: H = S + sum(p*p)/2,
  acceptance exp(Re(H0 - H1)), accept iff uniform <= acceptance.  Momenta and
  the uniform are injected (drawn by oracle.rng) instead of torch's global RNG.

Parity pins (tests/test_oracle_core.py): force against finite differences of
the action (Tests/Action/test_force.cpp:75-99), leapfrog reversibility and
O(eps^2) energy violation (Tests/Integrator/test_leapfrog.cpp), exact
arithmetic for the gauge pieces.
"""

import torch

from . import fermion as fm
from . import rng


class HubbardAction:
    def __init__(self, lat, Nt, beta, U, mu=0.0, stable_logdet=False, fermion_matrix=None):
        self.lat = lat
        self.Nt = Nt
        self.beta = beta
        self.U = U
        self.Utilde = (beta / Nt) * U                      # hubbard.tpp:10-19
        # HubbardFermionAction<..., FermionMatrixType> (hubbardFermiAction.tpp:23-60)
        self.hfm = (fermion_matrix or fm.HubbardExp)(lat, Nt, beta, mu)
        self.stable = stable_logdet

    def _logdet(self):
        return self.hfm.logdet_stable() if self.stable else self.hfm.logDetM()

    def gauge_eval(self, phi):
        return complex((phi * phi).sum()) / (2 * self.Utilde)

    def fermion_eval(self, phi):
        self.hfm.populate(phi, fm.PARTICLE)
        ld = self._logdet()
        self.hfm.populate(phi, fm.HOLE)
        ld += self._logdet()
        return -ld

    def eval(self, phi):
        phi = torch.as_tensor(phi, dtype=torch.complex128)
        return self.gauge_eval(phi) + self.fermion_eval(phi)

    def grad(self, phi):
        phi = torch.as_tensor(phi, dtype=torch.complex128)
        g = phi / self.Utilde
        self.hfm.populate(phi, fm.PARTICLE)
        g = g + self.hfm.gradLogDetM()
        self.hfm.populate(phi, fm.HOLE)
        g = g - self.hfm.gradLogDetM()
        return g

    def force(self, phi):
        return -self.grad(phi)


def leapfrog_real_force(action, q, p, traj_length, nmd):
    # leapfrogRealForce.tpp:35-74
    eps = traj_length / nmd
    q = torch.as_tensor(q, dtype=torch.complex128).clone()
    p = torch.as_tensor(p, dtype=torch.complex128).clone()

    def rforce(x):
        f = action.force(x)
        return torch.complex(f.real, torch.zeros_like(f.real))

    p = p + (0.5 * eps) * rforce(q)
    q = q + eps * p
    for _ in range(nmd - 1):
        p = p + eps * rforce(q)
        q = q + eps * p
    p = p + (0.5 * eps) * rforce(q)
    return q, p


def hmc_trajectory(action, phi, S0, momentum, uniform, traj_length, nmd):
    """HMC.tpp:185-241 with injected momentum (real, stored complex) and
    Metropolis uniform.  Returns dict(phi, S, dH, acc_prob, accepted)."""
    phi = torch.as_tensor(phi, dtype=torch.complex128)
    p = torch.as_tensor(momentum, dtype=torch.complex128)
    q1, p1 = leapfrog_real_force(action, phi, p, traj_length, nmd)
    S1 = action.eval(q1)
    H0 = S0 + 0.5 * complex((p * p).sum())
    H1 = S1 + 0.5 * complex((p1 * p1).sum())
    acc = float(torch.exp(torch.tensor((H0 - H1).real)))
    accepted = uniform <= acc
    return {"phi": q1 if accepted else phi, "S": S1 if accepted else S0,
            "dH": (H1 - H0), "acc_prob": acc, "accepted": accepted}


def momentum_draw(seed, traj, shape):
    """Momenta for trajectory `traj`: real N(0,1) (HMC.tpp:189-195 draws
    complex randn and zeroes Im; factory.tpp:38-41 gives Re ~ N(0,1))."""
    import numpy as np
    n = int(np.prod(shape))
    z = rng.normal(seed, 2 * traj, n).reshape(shape)
    return torch.as_tensor(z.astype(np.complex128))


def metropolis_uniform(seed, traj):
    return float(rng.uniform(seed, 2 * traj + 1, 1)[0])


def hmc_chain(action, phi0, ntraj, traj_length, nmd, seed):
    """ntraj trajectories from phi0 with momenta/uniforms from oracle.rng."""
    phi = torch.as_tensor(phi0, dtype=torch.complex128)
    S = action.eval(phi)
    out = []
    for k in range(ntraj):
        p = momentum_draw(seed, k, tuple(phi.shape))
        u = metropolis_uniform(seed, k)
        res = hmc_trajectory(action, phi, S, p, u, traj_length, nmd)
        phi, S = res["phi"], res["S"]
        out.append(res)
    return out


def hmc_generate(action, phi0, nconf, traj_length, nmd, seed, save_frequency=1, markov_time=0, all_states=True):
    """HMC::generate<AllStates | LastState>(state, Nconf, saveFrequency)
    (src/NSL/MarkovChain/HMC.tpp:57-155) without the HDF5 part, with
    MarkovState records (markovState.tpp:13-96) as dicts.  The trajectory
    that makes markov time m + 1 draws momentum_draw(seed, m) and
    metropolis_uniform(seed, m).  AllStates: states[0] = the start state,
    states[n] = saveFrequency trajectories after states[n-1] (:98-137);
    LastState: Nconf * saveFrequency - 1 trajectories (:144-152)."""
    save = save_frequency if save_frequency > 0 else 1          # :61-63
    phi = torch.as_tensor(phi0, dtype=torch.complex128)
    state = {"phi": phi, "S": action.eval(phi), "acc_prob": 1.0, "markov_time": markov_time,
             "accepted": True, "dH": 0j}

    def step(st):
        m = st["markov_time"]
        p = momentum_draw(seed, m, tuple(st["phi"].shape))
        u = metropolis_uniform(seed, m)
        res = hmc_trajectory(action, st["phi"], st["S"], p, u, traj_length, nmd)
        return {"phi": res["phi"], "S": res["S"], "acc_prob": res["acc_prob"], "markov_time": m + 1,
                "accepted": res["accepted"], "dH": res["dH"]}

    if all_states:
        MC = [state]
        for _ in range(1, nconf):
            tmp = MC[-1]
            for _ in range(save):
                tmp = step(tmp)
            MC.append(tmp)
        return MC
    for _ in range(1, nconf * save):
        state = step(state)
    return [state]
