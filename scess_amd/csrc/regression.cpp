// SCESS::Tests::RunRegressionTests: the closed-shell CC rows of the reference's regression runner
// (src/SCESS/Tests.cpp:494-737, run by `SCESS test`, Application.cpp:49-58), through this library's
// method API: each check prints "[PASS]" or "[FAIL]" with the value, the reference value, |diff| and
// the tolerance (CheckEnergy, Tests.cpp:77-91); the return value is the number of failed checks.
// The reference energies are the PySCF 2.14 values of Tests.cpp:32-55 for the built-in geometries.
// Rows outside the CC path (FCI, CISD, UHF, LDA, F12, the libint2-vs-CGF integrals and the basis-set
// library names) have no counterpart here.
#include "../../include/scess_cc.h"

#include <cmath>
#include <cstdio>
#include <exception>
#include <sstream>
#include <string>

namespace SCESS {
namespace Tests {
namespace {

namespace Reference {  // Tests.cpp:32-55
constexpr double H2_RHF = -1.1175058851;
constexpr double H2_MP2 = -0.0125038087;
constexpr double H2_CCSD = -0.0193419373;
constexpr double H2O_RHF = -74.9629832288;
constexpr double H2O_MP2 = -0.0355225667;
constexpr double H2O_CCSD = -0.0494026324;
constexpr double H2O_T = -0.0000673499;
constexpr double H2O_DZ_RHF = -76.0267859418;
constexpr double H2O_DZ_MP2 = -0.2039840885;
constexpr double H2O_DZ_CCSD = -0.2133077024;
constexpr double H2O_DZ_T = -0.0030574132;
constexpr double H2O_631G_RHF = -75.9839769556;
}  // namespace Reference

int gFailures = 0;

bool CheckEnergy(const char *label, double value, double reference, double tolerance) {
    const double diff = std::fabs(value - reference);
    const bool pass = diff <= tolerance;  // a NaN value fails
    std::printf("  [%s] %s: %.10f (reference %.10f, |diff| = %.2e, tolerance %.0e)\n", pass ? "PASS" : "FAIL",
                label, value, reference, diff, tolerance);
    std::fflush(stdout);
    if (!pass)
        gFailures++;
    return pass;
}

void Section(const char *title) {
    std::printf("\n--- %s ---\n", title);
    std::fflush(stdout);
}

struct System {
    Molecule mol;
    Integrals::MolInts mints;
    HF::HF_Wfn rhf;
};

System RunRHF(const Molecule &mol, const char *basisName) {
    System s;
    s.mol = mol;
    const Basis::BasisSet basis = Basis::LoadBasisSet(basisName);
    std::stringstream ss;
    s.rhf = HF::RHF(mol, basis, 1e-10, 100, &s.mints, nullptr, &ss);
    return s;
}

// A section whose setup throws (a missing basis set file, no GPU) counts as one failure.
template <class F>
void Guarded(const char *title, F body) {
    Section(title);
    try {
        body();
    } catch (const std::exception &ex) {
        std::printf("  [FAIL] %s: %s\n", title, ex.what());
        gFailures++;
    }
}

}  // namespace

int RunRegressionTests() {
    gFailures = 0;

    Guarded("H2 / STO-3G", [] {
        System s = RunRHF(Molecule::GenerateHydrogenMolecule(), "sto-3g");
        CheckEnergy("E(RHF)", s.rhf.E_HF, Reference::H2_RHF, 1e-8);
        CheckEnergy("E(MP2 correlation)", HF::MP2_RHF_ClosedShellAO(s.rhf, s.mints.eri), Reference::H2_MP2, 1e-8);
        CheckEnergy("E(CCSD correlation)", CC::CCSD_DIIS(s.rhf, s.mints).E_CCSD, Reference::H2_CCSD, 1e-6);
    });

    Guarded("H2O / STO-3G", [] {
        System s = RunRHF(Molecule::GenerateWaterMolecule(), "sto-3g");
        CheckEnergy("E(RHF)", s.rhf.E_HF, Reference::H2O_RHF, 1e-6);
        CheckEnergy("E(MP2 correlation)", HF::MP2_RHF_ClosedShellAO(s.rhf, s.mints.eri), Reference::H2O_MP2, 1e-7);
        std::stringstream ss;
        const CC::RCCSD_Wfn ccsd = CC::RCCSD_T(s.rhf, s.mints, nullptr, &ss);
        CheckEnergy("E(CCSD correlation)", ccsd.E_CCSD, Reference::H2O_CCSD, 1e-7);
        CheckEnergy("E(T)", ccsd.E_T, Reference::H2O_T, 1e-8);
        // the independent spin-orbital implementation on the same integrals and orbitals
        const CC::CCSD_Wfn spin = CC::CCSD_T(s.rhf, s.mints);
        CheckEnergy("E(T), closed-shell vs spin orbitals", ccsd.E_T, spin.E_T, 1e-7);
    });

    Guarded("H2O / cc-pVDZ (solid harmonic d functions)", [] {
        System s = RunRHF(Molecule::GenerateWaterMolecule(), "cc-pVDZ");
        CheckEnergy("E(RHF)", s.rhf.E_HF, Reference::H2O_DZ_RHF, 1e-8);
        CheckEnergy("E(MP2 correlation)", HF::MP2_RHF_ClosedShellAO(s.rhf, s.mints.eri), Reference::H2O_DZ_MP2,
                    1e-8);
        std::stringstream ss;
        const CC::RCCSD_Wfn ccsd = CC::RCCSD_T(s.rhf, s.mints, nullptr, &ss);
        CheckEnergy("E(CCSD correlation)", ccsd.E_CCSD, Reference::H2O_DZ_CCSD, 1e-7);
        CheckEnergy("E(T)", ccsd.E_T, Reference::H2O_DZ_T, 1e-7);
    });

    Guarded("H2O / 6-31G", [] {
        System s = RunRHF(Molecule::GenerateWaterMolecule(), "6-31G");
        CheckEnergy("E(RHF)", s.rhf.E_HF, Reference::H2O_631G_RHF, 1e-8);
    });

    Guarded("Closed-shell (spatial orbital) CCSD", [] {
        System h2 = RunRHF(Molecule::GenerateHydrogenMolecule(), "sto-3g");
        std::stringstream ss;
        CheckEnergy("H2 E(CCSD correlation)", CC::RCCSD_DIIS(h2.rhf, h2.mints, nullptr, &ss).E_CCSD,
                    Reference::H2_CCSD, 1e-8);
        System w = RunRHF(Molecule::GenerateWaterMolecule(), "sto-3g");
        const double E_RCCSD = CC::RCCSD_DIIS(w.rhf, w.mints, nullptr, &ss).E_CCSD;
        CheckEnergy("H2O E(CCSD correlation)", E_RCCSD, Reference::H2O_CCSD, 1e-7);
        CheckEnergy("H2O E(CCSD), spatial vs spin orbitals", E_RCCSD, CC::CCSD_DIIS(w.rhf, w.mints).E_CCSD, 1e-6);
    });

    if (gFailures == 0)
        std::printf("\nAll regression tests passed\n");
    else
        std::printf("\n%d regression test(s) failed\n", gFailures);
    std::fflush(stdout);
    return gFailures;
}

}  // namespace Tests
}  // namespace SCESS
