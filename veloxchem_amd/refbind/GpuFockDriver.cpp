// Reference-side drop-in of the GPU Fock build (see GpuFockDriver.hpp).
//
// make_shell_table recovers the basis from the screener alone, as
// CFockDriver::compute receives it (src/t4c_func/FockDriver.cpp:120-171):
// every contracted shell s appears in its own diagonal pair (s, s) of a
// same-block pair block (src/orbdata/GtoPairBlock.cpp:88-166), which holds its
// exponents, its centre, its AO index (component 0; index 0 of the orbital
// index arrays is the number of functions of that L, the component stride,
// src/orbdata/MolecularBasis.cpp:452-470) and the products N_k N_l of its
// normalised coefficients; N_0 = sqrt(N_0 N_0) and N_k = (N_k N_0) / N_0.
// This fixes each shell up to the sign of N_0.  The products N_a0 N_b0 of the
// off-diagonal pairs give the relative signs of connected shells; a sign
// common to a connected set of shells cancels in every integral, because a
// screened pair never joins two such sets.
#include "GpuFockDriver.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <queue>
#include <stdexcept>
#include <tuple>

#include "../../include/vlxfock.h"
#include "BlockedGtoPairBlock.hpp"
#include "GtoPairBlock.hpp"
#include "SubMatrix.hpp"

namespace {

void check(int rc)
{
    if (rc != VLX_OK) throw std::runtime_error(std::string("libvlxfock: ") + vlx_last_error());
}

}  // namespace

auto make_shell_table(const CT4CScreener& screener) -> CGpuShellTable
{
    struct Shell {
        int l = 0, atom = 0;
        size_t ao = 0;  // AO index of component 0
        std::array<double, 3> c{};
        std::vector<double> e, n;
    };
    std::map<std::pair<int, size_t>, Shell> shells;  // (l, AO index of component 0) -> shell
    std::map<int, size_t> n_l;                        // functions per angular momentum
    struct Link {
        std::pair<int, size_t> a, b;
        double prod;  // N_a0 N_b0 with the reference's signs
    };
    std::vector<Link> links;
    for (const auto& bpb : screener.gto_pair_blocks()) {
        for (int bin = 0; bin < 16; bin++) {
            if (bpb.is_empty_gto_pair_block(bin)) continue;
            const auto pb = bpb.gto_pair_block(bin);
            const auto [la, lb] = pb.angular_momentums();
            const size_t ncp = pb.number_of_contracted_pairs();
            const int npp = pb.number_of_primitive_pairs();
            const auto bidx = pb.bra_orbital_indices(), kidx = pb.ket_orbital_indices();
            n_l[la] = bidx[0];
            n_l[lb] = kidx[0];
            const auto bexp = pb.bra_exponents();
            const auto norms = pb.normalization_factors();
            const auto bc = pb.bra_coordinates();
            const auto ba = pb.bra_atomic_indices();
            for (size_t ij = 0; ij < ncp; ij++) {
                const auto ka = std::make_pair(la, bidx[ij + 1]), kb = std::make_pair(lb, kidx[ij + 1]);
                if (ka != kb) {
                    links.push_back({ka, kb, norms[ij]});
                    continue;
                }
                // diagonal pair: both sides are one shell of np primitives, npp = np^2
                const int np = static_cast<int>(std::lround(std::sqrt(static_cast<double>(npp))));
                if (np * np != npp) throw std::runtime_error("make_shell_table: diagonal pair with uneven primitives");
                if (shells.count(ka)) continue;
                Shell sh;
                sh.l = la;
                sh.atom = ba[ij];
                sh.ao = bidx[ij + 1];
                const auto xyz = bc[ij].coordinates();
                sh.c = {xyz[0], xyz[1], xyz[2]};
                const double p00 = norms[ij];
                if (!(p00 > 0.0)) throw std::runtime_error("make_shell_table: non-positive N_0^2");
                const double n0 = std::sqrt(p00);
                for (int k = 0; k < np; k++) {
                    const size_t k0 = (static_cast<size_t>(k) * np) * ncp + ij;  // primitive pair (k, 0)
                    sh.e.push_back(bexp[k0]);
                    sh.n.push_back(k == 0 ? n0 : norms[k0] / n0);
                }
                shells.emplace(ka, std::move(sh));
            }
        }
    }
    // angular block offsets of the AO order (L-major, then component, then function); every block
    // complete: component-0 AO indices off_l .. off_l + n_l - 1
    std::map<int, int> off_l;
    int nao = 0;
    for (const auto& [l, nf] : n_l) {
        off_l[l] = nao;
        nao += (2 * l + 1) * static_cast<int>(nf);
    }
    for (const auto& [l, nf] : n_l)
        for (size_t f = 0; f < nf; f++)
            if (!shells.count({l, static_cast<size_t>(off_l[l]) + f}))
                throw std::runtime_error("make_shell_table: the screener's pair blocks miss a diagonal pair");
    // relative signs along the off-diagonal pairs (breadth first from each unvisited shell)
    std::map<std::pair<int, size_t>, std::vector<std::pair<std::pair<int, size_t>, double>>> adj;
    for (const auto& lk : links) {
        adj[lk.a].push_back({lk.b, lk.prod});
        adj[lk.b].push_back({lk.a, lk.prod});
    }
    std::map<std::pair<int, size_t>, double> sign;
    for (const auto& kv : shells) {
        if (sign.count(kv.first)) continue;
        sign[kv.first] = 1.0;
        std::queue<std::pair<int, size_t>> q;
        q.push(kv.first);
        while (!q.empty()) {
            const auto u = q.front();
            q.pop();
            for (const auto& [v, prod] : adj[u]) {
                if (prod == 0.0) continue;
                // prod = N_u0 N_v0; ours are |N_u0| sign_u and |N_v0| sign_v
                const double want = (prod > 0.0 ? 1.0 : -1.0) * sign[u];
                const auto it = sign.find(v);
                if (it == sign.end()) {
                    sign[v] = want;
                    q.push(v);
                } else if (it->second != want) {
                    throw std::runtime_error("make_shell_table: inconsistent normalisation signs");
                }
            }
        }
    }
    // shells in atom order, then by (l, AO index)
    std::vector<const Shell*> order;
    for (const auto& kv : shells) order.push_back(&kv.second);
    std::stable_sort(order.begin(), order.end(), [](const Shell* x, const Shell* y) {
        return std::tie(x->atom, x->l, x->ao) < std::tie(y->atom, y->l, y->ao);
    });
    CGpuShellTable t;
    t.nao = nao;
    for (const Shell* sh : order) {
        const double sg = sign[{sh->l, sh->ao}];
        t.l.push_back(sh->l);
        t.nprim.push_back(static_cast<int>(sh->e.size()));
        t.prim_offset.push_back(static_cast<int>(t.exps.size()));
        t.exps.insert(t.exps.end(), sh->e.begin(), sh->e.end());
        for (double x : sh->n) t.norms.push_back(sg * x);
        t.centers.insert(t.centers.end(), sh->c.begin(), sh->c.end());
        t.ao_base.push_back(static_cast<int>(sh->ao));
        t.ao_stride.push_back(static_cast<int>(n_l[sh->l]));
        t.atom.push_back(sh->atom);
    }
    return t;
}

CGpuFockDriver::CGpuFockDriver(int device) : _device(device) { check(vlx_context_create(device, &_ctx)); }

CGpuFockDriver::~CGpuFockDriver()
{
    for (auto& kv : _screeners) vlx_screener_destroy(kv.second.second);
    vlx_context_destroy(_ctx);
}

auto CGpuFockDriver::set_block_size_factor(const int factor) -> void
{
    check(vlx_set_block_size_factor(_ctx, factor));
    _block_size_factor = factor;
}

auto CGpuFockDriver::_screener_for(const CT4CScreener& screener) const -> vlx_screener*
{
    const auto blocks = screener.gto_pair_blocks();
    auto it = _screeners.find(&screener);
    if (it != _screeners.end()) {
        if (it->second.first == blocks) return it->second.second;
        vlx_screener_destroy(it->second.second);
        _screeners.erase(it);
    }
    const CGpuShellTable t = make_shell_table(screener);
    vlx_screener* s = nullptr;
    check(vlx_screener_create(_ctx, static_cast<int>(t.l.size()), t.l.data(), t.nprim.data(), t.prim_offset.data(),
                              t.exps.data(), t.norms.data(), t.centers.data(), t.ao_base.data(), t.ao_stride.data(),
                              t.nao, &s));
    _screeners.emplace(&screener, std::make_pair(blocks, s));
    return s;
}

auto CGpuFockDriver::compute(const CT4CScreener& screener, const CMatrix& density, const std::string& label,
                             const double exchange_factor, const double omega, const int ithreshold) const -> CMatrix
{
    vlx_screener* s = _screener_for(screener);
    const auto full = density.full_matrix();
    const size_t n = full.number_of_rows();
    if (full.number_of_columns() != n) throw std::runtime_error("CGpuFockDriver: density is not square");
    std::vector<double> g(n * n);
    // symmetric densities take the symmetric build; general and antisymmetric ones the general
    // build (J of the symmetric part, K of both parts)
    if (density.get_type() == mat_t::symmetric)
        check(vlx_fock_compute(_ctx, s, full.data(), label.c_str(), exchange_factor, omega, ithreshold, g.data()));
    else
        check(vlx_fock_compute_general(_ctx, s, full.data(), label.c_str(), exchange_factor, omega, ithreshold,
                                       g.data()));
    // the result has the density's layout and type (src/t4c_func/FockDriver.cpp:131-133)
    CMatrix fock(density);
    fock.zero();
    for (const auto& key : fock.angular_pairs()) {
        auto* sub = fock.sub_matrix(key);
        const auto dims = sub->get_dimensions();
        for (size_t i = 0; i < dims[2]; i++)
            for (size_t j = 0; j < dims[3]; j++) sub->at({i, j}) = g[(i + dims[0]) * n + j + dims[1]];
    }
    return fock;
}

auto CGpuFockDriver::compute(const CT4CScreener& screener, const int rank, const int nodes, const CMatrix& density,
                             const std::string& label, const double exchange_factor, const double omega,
                             const int ithreshold) const -> CMatrix
{
    if (rank != 0 || nodes != 1)
        throw std::runtime_error("CGpuFockDriver: one GPU builds the whole matrix (rank 0 of 1 node only)");
    return compute(screener, density, label, exchange_factor, omega, ithreshold);
}

auto CGpuFockDriver::compute(const CT4CScreener& screener, const CMatrices& densities,
                             const std::vector<std::string>& labels, const double exchange_factor, const double omega,
                             const int ithreshold) const -> CMatrices
{
    const auto keys = densities.keys();
    if (keys.size() != labels.size()) throw std::runtime_error("CGpuFockDriver: one label per density");
    if (keys.empty()) return CMatrices(densities);
    vlx_screener* s = _screener_for(screener);
    // one pass over the integrals digests every density (vlx_fock_compute_multi); erf(omega r)/r only
    // when every label is range-separated (CT4CMatricesDistributor::need_omega,
    // src/t4c_func/T4CMatricesDistributor.cpp:55-66), as the library applies it
    const size_t k = keys.size();
    size_t n = 0;
    std::vector<double> dens, g;
    std::vector<int> general(k);
    std::vector<const char*> labs(k);
    for (size_t i = 0; i < k; i++) {
        const CMatrix* d = densities.matrix(keys[i]);
        const auto full = d->full_matrix();
        if (i == 0) {
            n = full.number_of_rows();
            dens.resize(k * n * n);
        }
        if (full.number_of_rows() != n || full.number_of_columns() != n)
            throw std::runtime_error("CGpuFockDriver: densities of different sizes");
        std::copy(full.data(), full.data() + n * n, dens.begin() + i * n * n);
        general[i] = d->get_type() == mat_t::symmetric ? 0 : 1;
        labs[i] = labels[i].c_str();
    }
    g.resize(k * n * n);
    check(vlx_fock_compute_multi(_ctx, s, static_cast<int>(k), dens.data(), general.data(), labs.data(),
                                 exchange_factor, omega, ithreshold, g.data()));
    CMatrices focks(densities);
    for (size_t i = 0; i < k; i++) {
        CMatrix* fock = focks.matrix(keys[i]);
        fock->zero();
        for (const auto& key : fock->angular_pairs()) {
            auto* sub = fock->sub_matrix(key);
            const auto dims = sub->get_dimensions();
            for (size_t r = 0; r < dims[2]; r++)
                for (size_t c = 0; c < dims[3]; c++) sub->at({r, c}) = g[i * n * n + (r + dims[0]) * n + c + dims[1]];
        }
    }
    return focks;
}
