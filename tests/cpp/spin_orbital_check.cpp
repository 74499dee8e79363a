// Host-only check of the spin-orbital CCSD(T) and of the host RHF path (include/scess_cc.h), run by
// tests/test_spin_orbital_cpu.py without a GPU: water / STO-3G at the reference's built-in geometry,
// host AO integrals and RHF, the MO transform <pq|rs> done here by plain loops, then
// SCESS::CC::CCSD_DIIS_FromMO with triples. Prints "E_RHF E_CCSD E_T" on one line.
#include "scess_cc.h"

#include <cstdio>
#include <exception>
#include <sstream>
#include <string>
#include <vector>

using namespace SCESS;

int main(int argc, char **argv) {
    try {
        const Molecule mol = argc > 1 && std::string(argv[1]) == "H2" ? Molecule::GenerateHydrogenMolecule()
                                                                      : Molecule::GenerateWaterMolecule();
        const Basis::BasisSet basis = Basis::LoadBasisSet("sto-3g");
        Integrals::MolInts mints;
        std::stringstream ss;
        const HF::HF_Wfn rhf = HF::RHF(mol, basis, 1e-10, 100, &mints, nullptr, &ss);
        const int64_t n = rhf.nmo, nao = rhf.nao;
        auto C = [&](int64_t mu, int64_t p) { return rhf.C_Alpha[(size_t)(mu + nao * p)]; };
        // (pq|rs) in four quarter transforms, then <pq|rs> = (pr|qs)
        std::vector<double> a((size_t)(nao * nao * nao * nao)), b(a.size());
        auto I = [&](int64_t i, int64_t j, int64_t k, int64_t l) { return (size_t)(((i * nao + j) * nao + k) * nao + l); };
        for (int64_t i = 0; i < nao; i++)
            for (int64_t j = 0; j < nao; j++)
                for (int64_t k = 0; k < nao; k++)
                    for (int64_t l = 0; l < nao; l++)
                        a[I(i, j, k, l)] = mints.eri(i, j, k, l);
        for (int pass = 0; pass < 4; pass++) {  // transform the leading index, rotate it to the back
            for (int64_t p = 0; p < n; p++)
                for (int64_t j = 0; j < nao; j++)
                    for (int64_t k = 0; k < nao; k++)
                        for (int64_t l = 0; l < nao; l++) {
                            double s = 0.0;
                            for (int64_t mu = 0; mu < nao; mu++)
                                s += C(mu, p) * a[I(mu, j, k, l)];
                            b[I(j, k, l, p)] = s;
                        }
            a.swap(b);
        }
        CC::TEITensor eri(n, n, n, n);
        for (int64_t p = 0; p < n; p++)
            for (int64_t q = 0; q < n; q++)
                for (int64_t r = 0; r < n; r++)
                    for (int64_t s = 0; s < n; s++)
                        eri(p, q, r, s) = a[I(p, r, q, s)];
        int64_t o = 0;
        while (o < (int64_t)rhf.Occ_Alpha.size() && rhf.Occ_Alpha[(size_t)o] == 1)
            o++;
        const CC::CCSD_Wfn w = CC::CCSD_DIIS_FromMO(eri, rhf.Ep_Alpha, o, true);
        std::printf("%.12f %.12f %.12f %d\n", rhf.E_HF, w.E_CCSD, w.E_T, w.converged ? 1 : 0);
        return w.converged ? 0 : 1;
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "spin_orbital_check: %s\n", ex.what());
        return 1;
    }
}
