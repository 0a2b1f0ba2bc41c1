// HubbardExp and HubbardDiag fermion matrices on the device (reference:
// src/NSL/FermionMatrix/Impl/hubbardExp.hpp:48-144, hubbardExp.tpp:12-125;
// hubbardDiag.hpp, hubbardDiag.tpp:10-165).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <tuple>

#include "engine.hpp"

namespace nslx {

int TileTable::nb_class(int64_t nb) {
    int c = 0;
    while ((int64_t(1) << c) < nb && c < 30) ++c;
    return c;
}

int TileTable::get(int prec, int nx, int nt, int64_t nb) {
    static const int forced = (int)env_double("NSLX_TILE", -1.0);
    if (forced >= 0 && forced < (int)tile_menu(prec).size() && tile_fits(prec, forced, nx)) return forced;
    int t;
    if (lookup(prec, nx, nt, nb, &t)) return t;
    return default_tile(prec, nx, nt, nb);
}

bool TileTable::lookup(int prec, int nx, int nt, int64_t nb, int* tile) {
    std::lock_guard<std::mutex> g(mu);
    auto it = best.find({prec, nx, nt, nb_class(nb)});
    if (it == best.end()) return false;
    *tile = it->second;
    return true;
}

void TileTable::set(int prec, int nx, int nt, int64_t nb, int tile) {
    std::lock_guard<std::mutex> g(mu);
    if (tile < 0) best.erase({prec, nx, nt, nb_class(nb)});
    else best[{prec, nx, nt, nb_class(nb)}] = tile;
}

void fm_init(FermionMatrix& m, Ctx* ctx, int nx, int nt, const double* K, double beta, double mu, int species,
             int disc) {
    NSLX_REQUIRE(K, "fm_create: null K");
    NSLX_REQUIRE(disc == DISC_EXP || disc == DISC_DIAG, "fm_create: bad discretisation");
    NSLX_REQUIRE(nx >= 1 && nt >= 2, "fm_create: need nx >= 1 and nt >= 2");
    NSLX_REQUIRE(species == NSLX_PARTICLE || species == NSLX_HOLE, "fm_create: bad species");
    NSLX_REQUIRE(std::isfinite(beta) && std::isfinite(mu), "fm_create: beta and mu must be finite");
    m.ctx = ctx;
    m.disc = disc;
    m.nx = nx;
    m.nt = nt;
    m.beta = beta;
    m.mu = mu;
    m.species = species;
    m.K_host.assign(K, K + (size_t)nx * nx);
    double nk = 0.0;
    for (int i = 0; i < nx; ++i) {
        double rs = 0.0;
        for (int j = 0; j < nx; ++j) rs += std::fabs(K[(size_t)i * nx + j]);
        nk = std::max(nk, rs);
    }
    m.normK = nk;
    m.K.alloc((size_t)nx * nx);
    NSLX_HIP(hipMemcpy(m.K.p, K, (size_t)nx * nx * sizeof(double), hipMemcpyHostToDevice));
}

double env_double(const char* name, double dflt) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    return std::atof(v);
}

EEntry& FermionMatrix::E(double a, hipStream_t s) {
    // memo keyed by the exponent, as SpatialLattice::exp_hopping_matrix
    // (src/NSL/Lattice/lattice.tpp:51-58)
    auto it = memo.find(a);
    if (it != memo.end()) return *it->second;
    auto e = std::make_unique<EEntry>();
    const int n = nx;
    e->dense.alloc((size_t)n * n);
    DevBuf<double> w1((size_t)n * n), w2((size_t)n * n);
    const double norm = std::fabs(a) * normK;
    int sq = 0;
    while (std::ldexp(norm, -sq) > 0.5) ++sq;
    expm_real_scaled(n, K.p, a, sq, e->dense.p, w1.p, w2.p, s);
    const size_t padded = (size_t)e_rows(n) * e_ld(n);
    e->e64.alloc(padded);
    e->e32.alloc(padded);
    e->ebf.alloc(2 * padded);
    pack_E(n, e->dense.p, e->e64.p, e->e32.p, e->ebf.p, s);
    NSLX_HIP(hipStreamSynchronize(s));   // w1/w2 are freed on return
    auto& ref = *e;
    memo.emplace(a, std::move(e));
    return ref;
}

EEntry& FermionMatrix::Elin(double a, double id, hipStream_t s) {
    // Lat.hopping_matrix(a) = a K (src/NSL/Lattice/lattice.tpp:46-49), plus id I
    auto it = memo_lin.find({a, id});
    if (it != memo_lin.end()) return *it->second;
    auto e = std::make_unique<EEntry>();
    const int n = nx;
    e->dense.alloc((size_t)n * n);
    scaled_plus_identity(n, K.p, a, id, e->dense.p, s);
    const size_t padded = (size_t)e_rows(n) * e_ld(n);
    e->e64.alloc(padded);
    e->e32.alloc(padded);
    e->ebf.alloc(2 * padded);
    pack_E(n, e->dense.p, e->e64.p, e->e32.p, e->ebf.p, s);
    auto& ref = *e;
    memo_lin.emplace(std::make_pair(a, id), std::move(e));
    return ref;
}

const void* FermionMatrix::Eptr(int prec, hipStream_t s) {
    EEntry& e = disc == DISC_DIAG ? Elin(sgn() * delta(), 0.0, s) : E(sgn() * delta(), s);
    if (prec == NSLX_F64) return e.e64.p;
    if (prec == NSLX_F32) return e.e32.p;
    return e.ebf.p;
}

void FermionMatrix::populate(const void* phi_dev, int64_t nf, int sp, hipStream_t s) {
    NSLX_REQUIRE(nf >= 1, "populate: nfield must be >= 1");
    NSLX_REQUIRE(phi_dev != nullptr, "populate: phi is null");
    if (sp >= 0) {
        NSLX_REQUIRE(sp == NSLX_PARTICLE || sp == NSLX_HOLE, "populate: bad species");
        species = sp;
    }
    const int64_t n = nf * len();
    if (nf != nfield) {
        phi.alloc(n);
        P64.alloc(n);
        Pinv64.alloc(n);
        P32.alloc(n);
        nfield = nf;
    }
    ++pop_gen;
    // deep copy of phi (the reference's populate copies into its own buffer)
    NSLX_HIP(hipMemcpyAsync(phi.p, phi_dev, n * sizeof(double2), hipMemcpyDeviceToDevice, s));
    phase_tables(phi.p, n, sgn(), delta() * mu, delta() * mu_im, P64.p, Pinv64.p, P32.p, s);
    // max |Im phi| bounds |P| (used by the solvers' attainable-accuracy test
    // and the time-slice chunk rule): reduced on the device, one value back
    mim_part.ensure(257);
    mim_h.ensure(1);
    max_abs_im(phi.p, n, mim_part.p, mim_part.p + 256, s);
    NSLX_HIP(hipMemcpyAsync(mim_h.p, mim_part.p + 256, sizeof(double), hipMemcpyDeviceToHost, s));
    NSLX_HIP(hipStreamSynchronize(s));
    max_im_phi = mim_h.p[0];
    // build (or find) the hop matrix of this species now, so that applies
    // after populate allocate nothing (nslx_fm_apply_noalloc relies on it)
    (void)Eptr(NSLX_F64, s);
}

double FermionMatrix::slice_growth() const {
    return std::fabs(delta()) * normK + mut_abs() + max_im_phi;
}

TimeSliceSolver& FermionMatrix::time_slice_solver(int prec, int nc, bool precond, hipStream_t s) {
    NSLX_REQUIRE(prec == NSLX_F64 || prec == NSLX_F32, "time-slice solve: precision must be F64 or F32");
    auto& t = tss[prec == NSLX_F64 ? 0 : 1];
    // a new factorisation when the fields changed (count, chunking or values)
    const bool wide = precond && prec == NSLX_F64;
    const int want = nc > 0 ? nc
                            : nt / tss_chunk(nt, slice_growth(), wide ? TSS_BOUND_PRECOND_F64 : TSS_BOUND_DIRECT,
                                             wide ? TSS_CAP_PRECOND : TSS_CAP_DIRECT);
    if (!t || t->chunks() != want || t->fields() != nfield) {
        t.reset();
        t = std::make_unique<TimeSliceSolver>(*this, prec, want, s);
    } else if (t->generation() != pop_gen) {
        t->refactor(s);
    }
    return *t;
}

void FermionMatrix::require_populated() const {
    if (nfield < 1) throw Error(NSLX_E_STATE, "fermion matrix used before populate()");
}

HopArgs FermionMatrix::hop_args(int prec, int64_t nb, hipStream_t s) {
    require_populated();
    NSLX_REQUIRE(nb >= 1, "nbatch must be >= 1");
    NSLX_REQUIRE(nb % nfield == 0, "nbatch must be a multiple of the number of populated fields");
    NSLX_REQUIRE(nb < (int64_t(1) << 31) / nt, "nbatch too large");
    HopArgs a;
    a.nx = nx;
    a.nt = nt;
    a.nb = (int)nb;
    a.rpf = (int)(nb / nfield);
    a.E = Eptr(prec, s);
    a.P = Pptr(prec);
    a.disc = disc;
    return a;
}

void FermionMatrix::hop(int prec, int op, const void* in, void* out, int64_t nb, hipStream_t s) {
    HopArgs a = hop_args(prec, nb, s);
    a.X = in;
    a.Y = out;
    launch_hop(prec, op, MODE_PLAIN, a, tile(prec, nb), s);
}

void FermionMatrix::apply(int op, const void* in, void* out, int64_t nb, int prec, hipStream_t s) {
    NSLX_REQUIRE(prec == NSLX_F64 || prec == NSLX_F32 || prec == NSLX_BF16X3, "apply: bad precision");
    NSLX_REQUIRE(in && out, "apply: null buffer");
    NSLX_REQUIRE(in != out, "apply: in and out must not overlap");
    switch (op) {
        case NSLX_OP_M: hop(prec, HOP_M, in, out, nb, s); break;
        case NSLX_OP_MDAGGER: hop(prec, HOP_MD, in, out, nb, s); break;
        case NSLX_OP_MMDAGGER:
        case NSLX_OP_MDAGGERM: {
            const size_t esz = prec == NSLX_F64 ? 16 : 8;
            work.ensure((size_t)nb * len() * esz);
            const int first = op == NSLX_OP_MMDAGGER ? HOP_MD : HOP_M;
            const int second = op == NSLX_OP_MMDAGGER ? HOP_M : HOP_MD;
            hop(prec, first, in, work.p, nb, s);
            hop(prec, second, work.p, out, nb, s);
            break;
        }
        default: throw Error(NSLX_E_ARG, "apply: bad op");
    }
}

}  // namespace nslx
