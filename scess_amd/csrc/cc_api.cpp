// SCESS::CC method API (include/scess_cc.h) over the C ABI (include/scess_gpu.h).
#include "../../include/scess_cc.h"
#include "../../include/scess_gpu.h"

#include <iomanip>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>

namespace SCESS::CC {
namespace {

struct Ctx {
    scess_ctx *c = nullptr;
    Ctx() {
        if (scess_create(&c, 0) != SCESS_OK)
            throw std::runtime_error("SCESS: no usable GPU (scess_create failed)");
    }
    ~Ctx() { scess_destroy(c); }
    void check(int rc) const {
        if (rc != SCESS_OK && rc != SCESS_CANCELLED)
            throw std::runtime_error(std::string("SCESS: ") + scess_last_error(c));
    }
};

int stop_cb(void *u) { return u && static_cast<const Thread *>(u)->ShouldStop() ? 1 : 0; }

// A ColMajor sub-block of <pq|rs>: blk(i0,i1,i2,i3) = eri(r[0]+i_{p0}, ...), with the block's
// index k reading eri's argument pos[k]. Returns the block in chemists' naming order.
template <class F>
std::vector<double> block(int64_t d0, int64_t d1, int64_t d2, int64_t d3, F f) {
    std::vector<double> out((size_t)(d0 * d1 * d2 * d3));
    size_t n = 0;
    for (int64_t l = 0; l < d3; l++)
        for (int64_t k = 0; k < d2; k++)
            for (int64_t j = 0; j < d1; j++)
                for (int64_t i = 0; i < d0; i++)
                    out[n++] = f(i, j, k, l);
    return out;
}

}  // namespace

RCCSD_Wfn run_ccsd_t(const TEITensor &eri, const std::vector<double> &e, int64_t nOcc, double E_HF,
                     const Thread *thread, std::stringstream *console);

RCCSD_Wfn SolveRCCSD_DIIS(const TEITensor &eri, const std::vector<double> &e, int64_t nOcc,
                          const RCCSD_F12Coupling *f12, const Thread *thread, std::stringstream *console) {
    const int64_t o = nOcc, n = eri.dimension(0), v = n - o;
    if (console)
        *console << std::setprecision(12)
                 << (f12 ? "--- Closed-shell CCSD with F12 geminals (CCSD-F12b) with DIIS acceleration ---\n"
                         : "--- Closed-shell Coupled Cluster Singles & Doubles (CCSD) with DIIS acceleration ---\n");
    // chemists' blocks from <pq|rs> = (pr|qs), as MakeIntegralBlocks (RCCSD.cpp:47-72)
    const auto ovov = block(o, v, o, v, [&](int64_t k, int64_t c, int64_t l, int64_t d) { return eri(k, l, o + c, o + d); });
    const auto oovv = block(o, o, v, v, [&](int64_t k, int64_t i, int64_t a, int64_t c) { return eri(k, o + a, i, o + c); });
    const auto ovvo = block(o, v, v, o, [&](int64_t k, int64_t c, int64_t a, int64_t i) { return eri(k, o + a, o + c, i); });
    const auto ovvv = block(o, v, v, v, [&](int64_t k, int64_t d, int64_t a, int64_t c) { return eri(k, o + a, o + d, o + c); });
    const auto ovoo = block(o, v, o, o, [&](int64_t l, int64_t c, int64_t k, int64_t i) { return eri(l, k, o + c, i); });
    const auto oooo = block(o, o, o, o, [&](int64_t k, int64_t i, int64_t l, int64_t j) { return eri(k, l, i, j); });
    const auto vvvv = block(v, v, v, v, [&](int64_t a, int64_t c, int64_t b, int64_t d) { return eri(o + a, o + b, o + c, o + d); });
    Ctx ctx;
    scess_ccsd_in in{};
    in.o = o;
    in.v = v;
    in.e = e.data();
    in.ovov = ovov.data();
    in.oovv = oovv.data();
    in.ovvo = ovvo.data();
    in.ovvv = ovvv.data();
    in.ovoo = ovoo.data();
    in.oooo = oooo.data();
    in.vvvv = vvvv.data();
    in.t2_driver = f12 ? f12->residual.data() : nullptr;
    in.f12_energy = f12 ? f12->energy.data() : nullptr;
    ctx.check(scess_ccsd_set_inputs(ctx.c, &in));
    ctx.check(scess_set_should_stop(ctx.c, stop_cb, (void *)thread));
    RCCSD_Wfn wfn;
    T1Tensor t1(o, v);
    T2Tensor t2(o, o, v, v);
    scess_ccsd_result res{};
    struct Log {
        std::stringstream *con;
    } lg{console};
    auto log = [](void *u, int it, double E, double dE, double rms) {
        auto *l = static_cast<Log *>(u);
        if (!l->con)
            return;
        if (it == 2)
            *l->con << "Turning on DIIS\n";
        *l->con << "   CCSD(" << it << "): E(corr) = " << E << "      dE = " << dE << "      rms(dt) = " << rms << '\n';
    };
    const int rc = scess_ccsd_solve(ctx.c, t1.data(), t2.data(), &res, log, &lg);
    if (rc == SCESS_CANCELLED)
        return {};
    ctx.check(rc);
    if (!res.converged) {
        if (console)
            *console << "!!! CCSD procedure failed to converge !!!\n\n\n";
        return wfn;
    }
    wfn.converged = true;
    wfn.E_CCSD = res.E_CCSD;
    wfn.E_F12_Coupling = res.E_F12_Coupling;
    wfn.t1 = std::move(t1);
    wfn.t2 = std::move(t2);
    if (console) {
        *console << "---- CCSD procedure converged! ----\n"
                 << "E(CCSD correlation) = " << wfn.E_CCSD << " hartrees\n";
        if (f12)
            *console << "   of which E(T2-geminal coupling) = " << wfn.E_F12_Coupling << " hartrees\n";
    }
    return wfn;
}

double RCCSD_TriplesCorrection(const TEITensor &eri, const std::vector<double> &e, int64_t nOcc, const T1Tensor &t1,
                               const T2Tensor &t2, const Thread *thread) {
    const int64_t o = nOcc, n = eri.dimension(0), v = n - o;
    // the elements of the repack of RCCSD_T.cpp:37-51
    const auto ovvv = block(o, v, v, v, [&](int64_t i, int64_t a, int64_t b, int64_t f) { return eri(i, o + b, o + a, o + f); });
    const auto ovoo = block(o, v, o, o, [&](int64_t i, int64_t a, int64_t m, int64_t j) { return eri(i, m, o + a, j); });
    const auto ovov = block(o, v, o, v, [&](int64_t i, int64_t a, int64_t j, int64_t b) { return eri(i, j, o + a, o + b); });
    Ctx ctx;
    const scess_t_in in{o, v, e.data(), t1.data(), t2.data(), ovvv.data(), ovoo.data(), ovov.data()};
    ctx.check(scess_set_should_stop(ctx.c, stop_cb, (void *)thread));
    ctx.check(scess_t_set_inputs(ctx.c, &in, 1 << SCESS_PREC_FP64));
    double E = 0.0;
    const int rc = scess_t_forward(ctx.c, SCESS_PREC_FP64, &E, nullptr);
    if (rc == SCESS_CANCELLED)
        return 0.0;
    ctx.check(rc);
    return E;
}

RCCSD_Wfn RCCSD_T_FromMO(const TEITensor &eri, const std::vector<double> &e, int64_t nOcc, double E_HF,
                         const Thread *thread, std::stringstream *console) {
    if (console)
        *console << "--- Closed-shell CCSD with perturbative triples (CCSD(T)) ---\n";
    return run_ccsd_t(eri, e, nOcc, E_HF, thread, console);
}

RCCSD_Wfn RCCSD_T(const HF::HF_Wfn &rhfWfn, const Integrals::MolInts &mints, const Thread *thread,
                  std::stringstream *console) {
    int64_t nOcc = 0;  // position of the first unoccupied orbital (RCCSD_T.cpp:172)
    while (nOcc < (int64_t)rhfWfn.Occ_Alpha.size() && rhfWfn.Occ_Alpha[nOcc] != 0)
        nOcc++;
    const TEITensor eri = Integrals::TransformTEI(mints.eri, rhfWfn.C_Alpha, rhfWfn.nmo, true);
    if (console)
        *console << "--- Closed-shell CCSD with perturbative triples (CCSD(T)) ---\n";
    return run_ccsd_t(eri, rhfWfn.Ep_Alpha, nOcc, rhfWfn.E_HF, thread, console);
}

RCCSD_Wfn RCCSD_DIIS(const HF::HF_Wfn &rhfWfn, const Integrals::MolInts &mints, const Thread *thread,
                     std::stringstream *console) {
    int64_t nOcc = 0;  // position of the first unoccupied orbital (RCCSD.cpp:398)
    while (nOcc < (int64_t)rhfWfn.Occ_Alpha.size() && rhfWfn.Occ_Alpha[nOcc] != 0)
        nOcc++;
    const TEITensor eri = Integrals::TransformTEI(mints.eri, rhfWfn.C_Alpha, rhfWfn.nmo, true);
    RCCSD_Wfn wfn = SolveRCCSD_DIIS(eri, rhfWfn.Ep_Alpha, nOcc, nullptr, thread, console);
    if (console && wfn.converged)
        *console << std::setprecision(12) << "E(Total) = E(Hartree Fock) + E(CCSD) = " << rhfWfn.E_HF + wfn.E_CCSD
                 << " hartrees\n\n\n";
    return wfn;
}

}  // namespace SCESS::CC

namespace SCESS::Integrals {

TEITensor TransformTEI(const TEITensor &tei, const std::vector<double> &C, int64_t nmo, bool convertChemToPhys) {
    const int64_t nao = tei.dimension(0);
    if ((int64_t)C.size() != nao * nmo)
        throw std::runtime_error("SCESS: TransformTEI: C is not nao x nmo");
    TEITensor out(nmo, nmo, nmo, nmo);
    scess_ctx *c = nullptr;
    if (scess_create(&c, 0) != SCESS_OK)
        throw std::runtime_error("SCESS: no usable GPU (scess_create failed)");
    const int rc = scess_transform_tei(c, nao, nmo, tei.data(), C.data(), convertChemToPhys ? 1 : 0, out.data());
    const std::string err = rc == SCESS_OK ? "" : scess_last_error(c);
    scess_destroy(c);
    if (rc != SCESS_OK)
        throw std::runtime_error("SCESS: " + err);
    return out;
}

}  // namespace SCESS::Integrals

namespace SCESS::CC {

RCCSD_Wfn run_ccsd_t(const TEITensor &eri, const std::vector<double> &e, int64_t nOcc, double E_HF,
                     const Thread *thread, std::stringstream *console) {
    RCCSD_Wfn wfn = SolveRCCSD_DIIS(eri, e, nOcc, nullptr, thread, console);
    if (!wfn.converged || (thread && thread->ShouldStop())) {
        if (console)
            *console << "Cannot calculate the triples correction as the CCSD amplitudes did not converge.\n";
        return wfn;
    }
    if (console)
        *console << "Calculating the triples correction to the energy...\n";
    wfn.E_T = RCCSD_TriplesCorrection(eri, e, nOcc, wfn.t1, wfn.t2, thread);
    if (console)
        *console << std::setprecision(12) << "E(T) = " << wfn.E_T << " hartrees\n"
                 << "E(Total) = E(Hartree Fock) + E(CCSD) + E((T)) = " << E_HF + wfn.E_CCSD + wfn.E_T
                 << " hartrees\n\n\n";
    return wfn;
}

}  // namespace SCESS::CC

namespace SCESS::CC {

struct TriplesFunction::State {
    scess_ctx *c = nullptr;
    int prec = SCESS_PREC_FP64;  // W recomputation of the backward
    int64_t o = 0, v = 0, n = 0;
    bool have = false;  // inputs of a successful forward are resident
    void check(int rc) const {
        if (rc != SCESS_OK && rc != SCESS_CANCELLED)
            throw std::runtime_error(std::string("SCESS: ") + scess_last_error(c));
    }
};

namespace {
int prec_code(TriplesFunction::Precision p) {
    switch (p) {
    case TriplesFunction::Precision::FP32: return SCESS_PREC_FP32;
    case TriplesFunction::Precision::BF16X3: return SCESS_PREC_BF16X3;
    case TriplesFunction::Precision::OZAKI: return SCESS_PREC_OZAKI;
    default: return SCESS_PREC_FP64;
    }
}
}  // namespace

TriplesFunction::TriplesFunction(Precision wRecompute, int device) : s_(new State) {
    s_->prec = prec_code(wRecompute);
    if (scess_create(&s_->c, device) != SCESS_OK) {
        delete s_;
        throw std::runtime_error("SCESS: no usable GPU (scess_create failed)");
    }
}

TriplesFunction::~TriplesFunction() {
    scess_destroy(s_->c);
    delete s_;
}

double TriplesFunction::forward(const TEITensor &eri, const std::vector<double> &e, int64_t nOcc, const T1Tensor &t1,
                                const T2Tensor &t2, const Thread *thread, Precision precision) {
    const int fprec = prec_code(precision);
    const int64_t o = nOcc, n = eri.dimension(0), v = n - o;
    s_->have = false;
    if (eri.dimension(1) != n || eri.dimension(2) != n || eri.dimension(3) != n || (int64_t)e.size() < n || o < 0 ||
        v < 0 || t1.dimension(0) != o || t1.dimension(1) != v || t2.dimension(0) != o || t2.dimension(1) != o ||
        t2.dimension(2) != v || t2.dimension(3) != v)
        throw std::runtime_error("SCESS: TriplesFunction::forward: inconsistent dimensions");
    // the elements of the repack of RCCSD_T.cpp:37-51, as RCCSD_TriplesCorrection
    const auto ovvv = block(o, v, v, v, [&](int64_t i, int64_t a, int64_t b, int64_t f) { return eri(i, o + b, o + a, o + f); });
    const auto ovoo = block(o, v, o, o, [&](int64_t i, int64_t a, int64_t m, int64_t j) { return eri(i, m, o + a, j); });
    const auto ovov = block(o, v, o, v, [&](int64_t i, int64_t a, int64_t j, int64_t b) { return eri(i, j, o + a, o + b); });
    const scess_t_in in{o, v, e.data(), t1.data(), t2.data(), ovvv.data(), ovoo.data(), ovov.data()};
    s_->check(scess_set_should_stop(s_->c, stop_cb, (void *)thread));
    s_->check(scess_t_set_inputs(s_->c, &in, (1 << SCESS_PREC_FP64) | (1 << s_->prec) | (1 << fprec)));
    double E = 0.0;
    const int rc = scess_t_forward(s_->c, fprec, &E, nullptr);
    (void)scess_set_should_stop(s_->c, nullptr, nullptr);  // the caller's thread may not outlive this call
    if (rc == SCESS_CANCELLED)
        return 0.0;
    s_->check(rc);
    s_->o = o;
    s_->v = v;
    s_->n = n;
    s_->have = true;
    return E;
}

TriplesGradient TriplesFunction::backward(double dE, const Thread *thread) {
    if (!s_->have)
        throw std::runtime_error("SCESS: TriplesFunction::backward needs a successful forward first");
    const int64_t o = s_->o, v = s_->v, n = s_->n;
    TriplesGradient g;
    g.ovvv = Tensor<4>(o, v, v, v);
    g.ovoo = Tensor<4>(o, v, o, o);
    g.ovov = Tensor<4>(o, v, o, v);
    g.t1 = T1Tensor(o, v);
    g.t2 = T2Tensor(o, o, v, v);
    scess_t_grad gr{g.ovvv.data(), g.ovoo.data(), g.ovov.data(), g.t1.data(), g.t2.data()};
    s_->check(scess_set_should_stop(s_->c, stop_cb, (void *)thread));
    const int rc = scess_t_backward(s_->c, s_->prec, dE, &gr);
    (void)scess_set_should_stop(s_->c, nullptr, nullptr);
    if (rc == SCESS_CANCELLED)
        return {};
    s_->check(rc);
    // each block element reads one entry of <pq|rs>, and the three index patterns are disjoint
    g.eri = TEITensor(n, n, n, n);
    for (int64_t f = 0; f < v; f++)
        for (int64_t b = 0; b < v; b++)
            for (int64_t a = 0; a < v; a++)
                for (int64_t i = 0; i < o; i++)
                    g.eri(i, o + b, o + a, o + f) = g.ovvv(i, a, b, f);
    for (int64_t j = 0; j < o; j++)
        for (int64_t m = 0; m < o; m++)
            for (int64_t a = 0; a < v; a++)
                for (int64_t i = 0; i < o; i++)
                    g.eri(i, m, o + a, j) = g.ovoo(i, a, m, j);
    for (int64_t b = 0; b < v; b++)
        for (int64_t j = 0; j < o; j++)
            for (int64_t a = 0; a < v; a++)
                for (int64_t i = 0; i < o; i++)
                    g.eri(i, j, o + a, o + b) = g.ovov(i, a, j, b);
    return g;
}

}  // namespace SCESS::CC

namespace SCESS::Driver {
namespace {

enum class Method { RHF, CCSD, CCSD_T };

bool run(Method method, const char *basisSetName, const Molecule &mol, const char *moleculeName) {
    const char *title = moleculeName ? moleculeName : "unknown";
    const char *what = method == Method::RHF ? "RHF" : method == Method::CCSD ? "CCSD" : "CCSD(T)";
    std::cout << "--- Performing " << what << " calculation on " << title << " ---\n" << std::setprecision(12);
    const Basis::BasisSet basis = Basis::LoadBasisSet(basisSetName ? basisSetName : "");
    Integrals::MolInts mints;
    std::stringstream console;
    const HF::HF_Wfn hf = HF::RHF(mol, basis, 1e-10, 100, method == Method::RHF ? nullptr : &mints, nullptr, &console);
    std::cout << console.str() << "\n";
    console.str("");
    if (method == Method::RHF) {
        std::cout << "\n\n--- Restricted Hartree-Fock Single Point Energy ---\n  Molecule = " << title
                  << "\n  E(RHF) = " << hf.E_HF << " hartrees\n";
        return true;
    }
    CC::RCCSD_Wfn cc;
    if (method == Method::CCSD_T) {
        cc = CC::RCCSD_T(hf, mints, nullptr, &console);
    } else {
        cc = CC::RCCSD_DIIS(hf, mints, nullptr, &console);
    }
    std::cout << console.str() << "\n";
    if (method == Method::CCSD_T)
        std::cout << "\n\n--- CCSD(T) Single Point Energy ---\n  Molecule = " << title << "\n  E(RHF) = " << hf.E_HF
                  << " hartrees\n  E(CCSD(T) correlation) = " << cc.E_CCSD + cc.E_T
                  << " hartrees\n  E(Total) = " << cc.E_CCSD + hf.E_HF + cc.E_T << " hartrees\n";
    else
        std::cout << "\n\n--- Coupled-Cluster Singles & Doubles Single Point Energy ---\n  Molecule = " << title
                  << "\n  E(RHF) = " << hf.E_HF << " hartrees\n  E(CCSD correlation) = " << cc.E_CCSD
                  << " hartrees\n  E(Total) = " << cc.E_CCSD + hf.E_HF << " hartrees\n";
    return cc.converged;
}

}  // namespace

bool RHF(const char *basisSetName, const Molecule &mol, const char *moleculeName) {
    return run(Method::RHF, basisSetName, mol, moleculeName);
}
bool CCSD(const char *basisSetName, const Molecule &mol, const char *moleculeName) {
    return run(Method::CCSD, basisSetName, mol, moleculeName);
}
bool CCSD_T(const char *basisSetName, const Molecule &mol, const char *moleculeName) {
    return run(Method::CCSD_T, basisSetName, mol, moleculeName);
}

}  // namespace SCESS::Driver
