"""HubbardExp and HubbardDiag fermion matrices, op for op on libtorch CPU
complex128 (oracle).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates src/NSL/FermionMatrix/Impl/hubbardExp.hpp:48-144 and
hubbardExp.tpp:12-211 with the tensor semantics of src/NSL/Tensor/Impl/
shift.tpp:59-71 (roll, then the wrapped slice times the boundary factor),
transpose.tpp (in-place transpose of the last two dims) and expand.tpp.
Fields have shape (..., Nt, Nx), x fastest.  s = +1 particle, -1 hole;
delta = beta/Nt; mu~ = delta*mu; P = exp(i s phi + s mu~), Pinv = exp(-i s phi
- s mu~) (hubbardExp.hpp:116-128).

Parity pins (tests/test_oracle_core.py): M, M^dagger against the dense
matrix built from the index formula of hubbardExp.hpp:22-31 on tiny lattices
(brute force), M^dagger = (M_dense)^H, MMdagger = M(Mdagger(.)), MdaggerM =
Mdagger(M(.)), Hermiticity, the phi = 0 Matsubara spectrum; logDetM against
LU of the dense matrix, the phi = 0 and uniform-slice closed forms, time-shift
and 2 pi invariance; gradLogDetM against central finite differences of
logDetM; logdet_stable against the same closed forms at large beta where the
reference form overflows.  HubbardDiag (hubbardDiag.tpp): M against the
brute-force index formula, M^dagger = (M_dense)^H, the reference's
term-by-term MMdagger / MdaggerM against M(Mdagger(.)) / Mdagger(M(.)),
logDetM against the LU of the dense matrix at mu = 0 (mod 2 pi i),
logdet_stable against logDetM, gradLogDetM against central differences.
"""

import math
import torch

PARTICLE = 0
HOLE = 1


def _shift(t, shift, dim, boundary=None):
    """Tensor::shift (shift.tpp:59-71): roll along dim, then multiply the
    wrapped slice by boundary (first `shift` slices for shift>0, last |shift|
    for shift<0)."""
    out = torch.roll(t, shifts=shift, dims=dim)
    if boundary is not None and shift != 0:
        out = out.clone()
        n = out.shape[dim]
        if shift > 0:
            idx = torch.arange(0, shift)
        else:
            idx = torch.arange(n + shift, n)
        sl = out.index_select(dim, idx) * boundary
        out.index_copy_(dim, idx, sl)
    return out


class HubbardExp:
    """Exponential discretisation  M_{tx,iy} = d_ti d_xy - [exp(s delta K)]_xy
    exp(i s phi_iy + s mu~) B_t d_{t,i+1},  B_0 = -1 (hubbardExp.hpp:22-31)."""

    def __init__(self, lat, Nt, beta=1.0, mu=0.0, species=PARTICLE):
        # hubbardExp.hpp:48-76
        self.Lat = lat
        self.Nt = Nt
        self.species = species
        self.delta = complex(beta) / Nt
        self.mu = (complex(beta) / Nt) * mu
        self.sgn = +1 if species == PARTICLE else -1
        self.phi = None
        self.phiExp = None
        self.phiExpInv = None

    def populate(self, phi, species=None):
        # hubbardExp.hpp:103-128
        if species is not None:
            self.species = species
            self.sgn = +1 if species == PARTICLE else -1
        phi = torch.as_tensor(phi, dtype=torch.complex128)
        self.phi = phi.clone()
        self.phiExp = torch.exp(complex(0, self.sgn) * phi + self.sgn * self.mu)
        self.phiExpInv = torch.exp(complex(0, -self.sgn) * phi - self.sgn * self.mu)

    def _E(self, a):
        return self.Lat.exp_hopping_matrix(a)

    def F_(self, psi):
        # hubbardExp.tpp:12-36
        Fpsi = torch.matmul(self._E(self.sgn * self.delta),
                            (self.phiExp * psi).transpose(-1, -2)).transpose(-1, -2)
        return _shift(Fpsi, 1, -2, -1.0)

    def M(self, psi):
        # hubbardExp.tpp:38-41
        psi = torch.as_tensor(psi, dtype=torch.complex128)
        return psi - self.F_(psi)

    def Mdagger(self, psi):
        # hubbardExp.tpp:43-75
        psi = torch.as_tensor(psi, dtype=torch.complex128)
        BexpKpsi = torch.matmul(self._E(self.sgn * self.delta.conjugate()),
                                psi.transpose(-1, -2).clone()).transpose(-1, -2)
        BexpKpsi = _shift(BexpKpsi, -1, -2, -1.0)
        return psi - torch.conj(self.phiExp) * BexpKpsi

    def MMdagger(self, psi):
        # hubbardExp.tpp:77-104 (the reference's 4-GEMM form)
        psi = torch.as_tensor(psi, dtype=torch.complex128)
        inner = torch.matmul(self._E(self.sgn * self.delta.conjugate()),
                             psi.transpose(-1, -2).clone()).transpose(-1, -2)
        absP = _shift(self.phiExp * torch.conj(self.phiExp), 1, -2)
        diag = torch.matmul(self._E(self.sgn * self.delta),
                            (absP * inner).transpose(-1, -2)).transpose(-1, -2)
        return (self.M(psi) + self.Mdagger(psi) - psi) + diag

    def MdaggerM(self, psi):
        # hubbardExp.tpp:106-125
        psi = torch.as_tensor(psi, dtype=torch.complex128)
        E2 = self._E(self.sgn * (self.delta.conjugate() + self.delta))
        blk = torch.matmul(E2, (self.phiExp * psi).transpose(-1, -2)).transpose(-1, -2)
        return self.M(psi) + self.Mdagger(psi) - psi + torch.conj(self.phiExp) * blk

    def logDetM(self):
        # hubbardExp.tpp:128-150: prod_t = E(s delta) diag(P_{t+1}),
        # sausage = prod_{Nt-1} ... prod_0, log(det(1 + sausage))
        Nt, Nx = self.phi.shape
        E = self._E(self.sgn * self.delta)
        Psh = torch.roll(self.phiExp, -1, 0)            # LinAlg::shift(P,-1)
        prod = E.unsqueeze(0) * Psh.unsqueeze(1)         # [t][a][b] = E_ab P'_{t,b}
        sausage = torch.eye(Nx, dtype=torch.complex128)
        for t in range(Nt - 1, -1, -1):
            sausage = torch.matmul(sausage, prod[t])
        return complex(torch.log(torch.linalg.det(torch.eye(Nx, dtype=torch.complex128)
                                                  + sausage)))

    def slice_matrices(self):
        """B_t = E(s delta) diag(P_{t+1}) in product order (for logdet_stable)."""
        E = self._E(self.sgn * self.delta)
        Psh = torch.roll(self.phiExp, -1, 0)
        return E.unsqueeze(0) * Psh.unsqueeze(1)

    def logdet_stable(self, nchunks=None):
        """log det M without forming the full slice product.

        Project addition for large beta, where the reference's logDetM
        overflows (det(1+prod) > e^709).  Uses det M = det(1 + B_{Nt-1}..B_0)
        = det of the block-cyclic matrix with blocks I on the diagonal,
        -C_j below it and +C_0 in the corner, for ANY grouping of consecutive
        slices into chunk products C_j; each C_j is well conditioned, and the
        reduced matrix is factorised by LU with partial pivoting
        (torch.linalg.slogdet).  Returns log|det| + i arg(det) with the
        imaginary part reduced to (-pi, pi].
        """
        Nt, Nx = self.phi.shape
        B = self.slice_matrices()
        if nchunks is None:
            nchunks = default_nchunks(self.Lat, self.delta, Nt, self.phi)
        nchunks = max(1, min(nchunks, Nt))
        bounds = [round(j * Nt / nchunks) for j in range(nchunks + 1)]
        C = []
        for j in range(nchunks):
            c = torch.eye(Nx, dtype=torch.complex128)
            for t in range(bounds[j + 1] - 1, bounds[j] - 1, -1):
                c = torch.matmul(c, B[t])
            C.append(c)
        n = nchunks * Nx
        R = torch.eye(n, dtype=torch.complex128)
        if nchunks == 1:
            R = R + C[0]
        else:
            R[0:Nx, (nchunks - 1) * Nx:n] += C[0]
            for j in range(1, nchunks):
                R[j * Nx:(j + 1) * Nx, (j - 1) * Nx:j * Nx] -= C[j]
        sign, logabs = torch.linalg.slogdet(R)
        ang = math.atan2(float(sign.imag), float(sign.real))
        return complex(float(logabs), ang)

    def gradLogDetM(self):
        # hubbardExp.tpp:152-211
        Nt, Nx = self.phi.shape
        II = 1j
        Pi = torch.roll(self.phiExpInv, 1, 0)             # LinAlg::shift(Pinv,+1)
        Fk = Pi.unsqueeze(2) * self._E(-1 * self.sgn * self.delta).unsqueeze(0)
        FkFkFk = torch.empty((Nt, Nx, Nx), dtype=torch.complex128)
        FkFkFk[Nt - 1] = Fk[Nt - 1]
        for t in range(Nt - 2, -1, -1):
            FkFkFk[t] = torch.matmul(Fk[t], FkFkFk[t + 1])
        invAp1F = torch.linalg.inv(torch.eye(Nx, dtype=torch.complex128) + FkFkFk[0])
        pi_dot = torch.empty((Nt, Nx), dtype=torch.complex128)
        pi_dot[Nt - 1] = II * torch.diagonal(torch.matmul(FkFkFk[0], invAp1F))
        for t in range(Nt - 1):
            invAp1F = torch.matmul(invAp1F, Fk[t])
            pi_dot[t] = II * torch.diagonal(torch.matmul(FkFkFk[t + 1], invAp1F))
        return pi_dot

    def M_dense(self):
        """Dense (Nt,Nx,Nt,Nx) matrix by applying M to unit vectors
        (fermionMatrix.hpp:87-110); column (i,y) = M e_{iy}."""
        Nt, Nx = self.phi.shape
        eye = torch.eye(Nt * Nx, dtype=torch.complex128).reshape(Nt * Nx, Nt, Nx)
        cols = self.M(eye)                               # batch of M e_k
        return cols.reshape(Nt * Nx, Nt * Nx).transpose(0, 1).reshape(Nt, Nx, Nt, Nx)


def default_nchunks(lat, delta, Nt, phi):
    """Number of slice groups for logdet_stable: keep each chunk product's
    growth below about e^8 (|E(delta)| grows like e^{|delta| ||K||}; |P| adds
    e^{|Im phi|})."""
    normK = float(torch.linalg.matrix_norm(lat.hops, ord=2))
    growth = abs(delta) * normK * Nt
    if phi is not None:
        growth += float(torch.abs(torch.as_tensor(phi).imag).sum(0).max())
    return max(1, int(math.ceil(growth / 8.0)))


def index_formula_dense(lat, Nt, beta, phi, mu=0.0, species=PARTICLE):
    """Dense M straight from the index formula of hubbardExp.hpp:22-31 (brute
    force, independent of the shift/roll code path above)."""
    Nx = lat.sites()
    s = +1 if species == PARTICLE else -1
    delta = complex(beta) / Nt
    E = lat.exp_hopping_matrix(s * delta)
    P = torch.exp(1j * s * torch.as_tensor(phi, dtype=torch.complex128) + s * delta * mu)
    D = torch.zeros((Nt, Nx, Nt, Nx), dtype=torch.complex128)
    for t in range(Nt):
        D[t, :, t, :] += torch.eye(Nx, dtype=torch.complex128)
        i = (t - 1) % Nt
        B = -1.0 if t == 0 else 1.0
        D[t, :, i, :] -= B * E * P[i].unsqueeze(0)
    return D


class HubbardDiag(HubbardExp):
    """Diagonal discretisation (src/NSL/FermionMatrix/Impl/hubbardDiag.hpp,
    hubbardDiag.tpp:10-252): M_{tx,iy} = d_ti d_xy - [s delta K]_xy d_ti
    - B_t exp(i s phi_iy + s mu~) d_xy d_{t,i+1} (the formula of the comment
    at hubbardDiag.tpp:13-18; the header's comment repeats the exponential
    form).  Constructor and populate are HubbardExp's (hubbardDiag.hpp).
    The reference rolls dimension 0 (Tensor::shift(n)); here the time axis
    is dimension -2, the same for the (Nt, Nx) fields the reference takes."""

    def _K(self, a):
        return self.Lat.hopping_matrix(a)

    def _Kmul(self, a, psi):
        return torch.matmul(self._K(a), psi.transpose(-1, -2)).transpose(-1, -2)

    def M(self, psi):
        # hubbardDiag.tpp:10-36
        psi = torch.as_tensor(psi, dtype=torch.complex128)
        out = _shift(self.phiExp * psi, 1, -2, -1.0)
        out = out + self._Kmul(self.sgn * self.delta, psi)
        return psi - out

    def Mdagger(self, psi):
        # hubbardDiag.tpp:39-72: copy, slice 0 times -1, shift(-1), conj(P) *
        psi = torch.as_tensor(psi, dtype=torch.complex128)
        B = torch.conj(self.phiExp) * _shift(psi, -1, -2, -1.0)
        return psi - self._Kmul(self.sgn * self.delta.conjugate(), psi) - B

    def MMdagger(self, psi):
        # hubbardDiag.tpp:75-125, term by term
        psi = torch.as_tensor(psi, dtype=torch.complex128)
        d, dc = self.sgn * self.delta, self.sgn * self.delta.conjugate()
        out = self.M(psi) + self.Mdagger(psi) - psi + self._Kmul(d, self._Kmul(dc, psi))
        out = out + self._Kmul(d, torch.conj(self.phiExp) * _shift(psi, -1, -2, -1.0))
        B = _shift(self.phiExp * self._Kmul(dc, psi), 1, -2, -1.0)
        return out + B + _shift(self.phiExp * torch.conj(self.phiExp), 1, -2) * psi

    def MdaggerM(self, psi):
        # hubbardDiag.tpp:128-164, term by term
        psi = torch.as_tensor(psi, dtype=torch.complex128)
        d, dc = self.sgn * self.delta, self.sgn * self.delta.conjugate()
        out = self.M(psi) + self.Mdagger(psi) - psi + self._Kmul(d, self._Kmul(dc, psi))
        out = out + torch.conj(self.phiExp) * self._Kmul(d, _shift(psi, -1, -2, -1.0))
        B = _shift(self._Kmul(dc, self.phiExp * psi), 1, -2, -1.0)
        return out + B + self.phiExp * torch.conj(self.phiExp) * psi

    def _D(self):
        Nx = self.phi.shape[-1]
        return torch.eye(Nx, dtype=torch.complex128) - self._K(self.sgn * self.delta)

    def logDetM(self):
        # hubbardDiag.tpp:166-190: fk_t = D diag(Pinv_{t+1}), D = 1 - s delta K,
        # sausage = fk_0 fk_1 ... fk_{Nt-1}; i s sum(phi) + log det(1 + sausage)
        Nt, Nx = self.phi.shape
        Pn = torch.roll(self.phiExpInv, -1, 0)          # LinAlg::shift(Pinv,-1)
        fk = self._D().unsqueeze(0) * Pn.unsqueeze(1)     # [t][k][x] = D_kx Pinv_{t+1,x}
        sausage = torch.eye(Nx, dtype=torch.complex128)
        for t in range(Nt):
            sausage = torch.matmul(sausage, fk[t])
        logdet = complex(torch.log(torch.linalg.det(torch.eye(Nx, dtype=torch.complex128) + sausage)))
        return 1j * self.sgn * complex(self.phi.sum()) + logdet

    def slice_matrices(self):
        """fk_t in product order for logdet_stable: det(1 + B_{Nt-1}..B_0)
        with B_t = fk_{Nt-1-t} is det(1 + fk_0 ... fk_{Nt-1})."""
        Pn = torch.roll(self.phiExpInv, -1, 0)
        fk = self._D().unsqueeze(0) * Pn.unsqueeze(1)
        return torch.flip(fk, dims=[0])

    def logdet_stable(self, nchunks=None):
        """logDetM through the block-cyclic LU (see HubbardExp.logdet_stable):
        log det(1 + sausage) reduced to (-pi, pi], then + i s sum(phi)."""
        ld = super().logdet_stable(nchunks)
        return ld + 1j * self.sgn * complex(self.phi.sum())

    def gradLogDetM(self):
        # hubbardDiag.tpp:193-252: Fk_t = diag(Pinv_{t-1}) D, then as HubbardExp
        Nt, Nx = self.phi.shape
        Pi = torch.roll(self.phiExpInv, 1, 0)             # LinAlg::shift(Pinv,+1)
        Fk = Pi.unsqueeze(2) * self._D().unsqueeze(0)
        FkFkFk = torch.empty((Nt, Nx, Nx), dtype=torch.complex128)
        FkFkFk[Nt - 1] = Fk[Nt - 1]
        for t in range(Nt - 2, -1, -1):
            FkFkFk[t] = torch.matmul(Fk[t], FkFkFk[t + 1])
        invAp1F = torch.linalg.inv(torch.eye(Nx, dtype=torch.complex128) + FkFkFk[0])
        pi_dot = torch.empty((Nt, Nx), dtype=torch.complex128)
        pi_dot[Nt - 1] = 1j * torch.diagonal(torch.matmul(FkFkFk[0], invAp1F))
        for t in range(Nt - 1):
            invAp1F = torch.matmul(invAp1F, Fk[t])
            pi_dot[t] = 1j * torch.diagonal(torch.matmul(FkFkFk[t + 1], invAp1F))
        return pi_dot


def diag_index_formula_dense(lat, Nt, beta, phi, mu=0.0, species=PARTICLE):
    """Dense HubbardDiag M from the index formula (hubbardDiag.tpp:13-18),
    brute force."""
    Nx = lat.sites()
    s = +1 if species == PARTICLE else -1
    delta = complex(beta) / Nt
    A = lat.hopping_matrix(s * delta)
    P = torch.exp(1j * s * torch.as_tensor(phi, dtype=torch.complex128) + s * delta * mu)
    D = torch.zeros((Nt, Nx, Nt, Nx), dtype=torch.complex128)
    for t in range(Nt):
        D[t, :, t, :] += torch.eye(Nx, dtype=torch.complex128) - A
        i = (t - 1) % Nt
        B = -1.0 if t == 0 else 1.0
        D[t, :, i, :] -= B * torch.diag(P[i])
    return D
