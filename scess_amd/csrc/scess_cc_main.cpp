// scess_cc: command-line runner of the closed-shell CC methods.
//
//   scess_cc rhf|ccsd|ccsd_t <molecule> <basis>   as the reference's `SCESS ccsd_t H2O cc-pVDZ`
//   scess_cc ccsd|ccsd_t|hf  <FCIDUMP>            from canonical RHF molecular orbitals
//   scess_cc test                                 the regression runner, as the reference's `SCESS test`
//
// The first form follows the reference's CLI (src/SCESS/Application.cpp:23-277) and drivers
// (Driver.cpp:51-65, 173-231): a built-in molecule (H2, LiH, NH3, H2O), a BSE-JSON basis set found
// by name under $SCESS_BASIS_DIR (the reference's Assets/Basis layout), AO integrals and RHF on the
// host, then the MO transform, CCSD and (T) on the GPU; `rhf` needs no GPU. The second form starts
// from an FCIDUMP (NORB, NELEC, one-electron h_pq, chemists' (pq|rs) with 8-fold symmetry, core
// energy) and prints the reference's console output of the CC driver.
// Orbital energies and E_HF follow from the dump: e_p = h_pp + sum_k [2(pp|kk) - (pk|kp)],
// E_HF = E_core + sum_k (h_kk + e_k). Exit code 0 on convergence, 1 otherwise (as the reference).
#include "../../include/scess_cc.h"

#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <string>

namespace {

struct Dump {
    int norb = 0, nelec = 0;
    double core = 0.0;
    std::vector<double> h;    // norb^2
    std::vector<double> eri;  // (pq|rs) chemists', norb^4, all 8 images filled
};

int header_int(const std::string &hdr, const std::string &key) {
    std::string up;
    for (char ch : hdr)
        up += (char)std::toupper((unsigned char)ch);
    const size_t p = up.find(key + "=");
    if (p == std::string::npos)
        throw std::runtime_error("FCIDUMP header has no " + key);
    return std::stoi(up.substr(p + key.size() + 1));
}

Dump read_fcidump(const std::string &path) {
    std::ifstream in(path);
    if (!in)
        throw std::runtime_error("cannot open " + path);
    std::string hdr, line;
    while (std::getline(in, line)) {
        hdr += line + " ";
        std::string up;
        for (char ch : line)
            up += (char)std::toupper((unsigned char)ch);
        if (up.find("&END") != std::string::npos || up.find('/') != std::string::npos)
            break;
    }
    Dump d;
    d.norb = header_int(hdr, "NORB");
    d.nelec = header_int(hdr, "NELEC");
    const int64_t n = d.norb;
    d.h.assign(n * n, 0.0);
    d.eri.assign(n * n * n * n, 0.0);
    auto E = [&](int64_t p, int64_t q, int64_t r, int64_t s) -> double & { return d.eri[((p * n + q) * n + r) * n + s]; };
    double x;
    int64_t i, j, k, l;
    while (in >> x >> i >> j >> k >> l) {
        if (i < 0 || j < 0 || k < 0 || l < 0 || i > n || j > n || k > n || l > n)
            throw std::runtime_error("FCIDUMP index out of range");
        if (i == 0 && j == 0 && k == 0 && l == 0) {
            d.core = x;
        } else if (i > 0 && j > 0 && k == 0 && l == 0) {
            d.h[(i - 1) * n + (j - 1)] = d.h[(j - 1) * n + (i - 1)] = x;
        } else if (i > 0 && j > 0 && k > 0 && l > 0) {
            const int64_t p = i - 1, q = j - 1, r = k - 1, s = l - 1;
            E(p, q, r, s) = E(q, p, r, s) = E(p, q, s, r) = E(q, p, s, r) = x;
            E(r, s, p, q) = E(s, r, p, q) = E(r, s, q, p) = E(s, r, q, p) = x;
        } else if (i > 0 && j == 0 && k == 0 && l == 0) {
            // (i 0 0 0): an orbital energy, written by some programs; recomputed from h and (pq|rs) instead
        } else {
            throw std::runtime_error("FCIDUMP line with an invalid index pattern: " + std::to_string(i) + " " +
                                     std::to_string(j) + " " + std::to_string(k) + " " + std::to_string(l));
        }
    }
    return d;
}

// Driver.cpp:51-65 (rhf), 173-200 (ccsd), 203-231 (ccsd_t) with a built-in molecule and a basis set
int run_molecule(const std::string &method, const std::string &molName, const std::string &basisName) {
    // the names the reference's main passes to its drivers (Application.cpp:176-194)
    static const std::map<std::string, std::string> longName = {
        {"H2", "Hydrogen (H2)"}, {"LiH", "Lithium Hydride (LiH)"}, {"H2O", "Water (H2O)"}, {"NH3", "Ammonia (NH3)"}};
    const SCESS::Molecule mol = SCESS::Molecule::FromName(molName);
    const std::string title = longName.at(molName);
    const bool ok = method == "rhf"    ? SCESS::Driver::RHF(basisName.c_str(), mol, title.c_str())
                    : method == "ccsd" ? SCESS::Driver::CCSD(basisName.c_str(), mol, title.c_str())
                                       : SCESS::Driver::CCSD_T(basisName.c_str(), mol, title.c_str());
    return ok ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc == 2 && std::string(argv[1]) == "test") {
        // the reference's `SCESS test` (Application.cpp:49-58): exit code 0 only when every check passes
        const int failures = SCESS::Tests::RunRegressionTests();
        return failures == 0 ? 0 : 1;
    }
    if (argc == 4) {
        std::string method = argv[1];
        for (char &c : method)
            c = (char)std::tolower((unsigned char)c);
        if (method != "rhf" && method != "ccsd" && method != "ccsd_t") {
            std::fprintf(stderr, "%s is not a valid method (rhf, ccsd, ccsd_t)\n", argv[1]);
            return 1;
        }
        try {
            return run_molecule(method, argv[2], argv[3]);
        } catch (const std::exception &ex) {
            std::fprintf(stderr, "scess_cc: %s\n", ex.what());
            return 1;
        }
    }
    const std::string method = argc == 3 ? argv[1] : "";
    if (method != "ccsd" && method != "ccsd_t" && method != "hf") {
        std::fprintf(stderr,
                     "usage: %s rhf|ccsd|ccsd_t <molecule> <basis>\n       %s ccsd|ccsd_t|hf <FCIDUMP>\n       %s test\n",
                     argv[0], argv[0], argv[0]);
        return 1;
    }
    const bool triples = method == "ccsd_t";
    try {
        const Dump d = read_fcidump(argv[2]);
        if (d.nelec % 2 != 0)
            throw std::runtime_error("closed-shell methods need an even number of electrons");
        const int64_t n = d.norb, o = d.nelec / 2;
        auto G = [&](int64_t p, int64_t q, int64_t r, int64_t s) { return d.eri[((p * n + q) * n + r) * n + s]; };
        std::vector<double> e(n);
        double E_HF = d.core, offdiag = 0.0;
        for (int64_t p = 0; p < n; p++)
            for (int64_t q = 0; q < n; q++) {
                double f = d.h[p * n + q];
                for (int64_t k = 0; k < o; k++)
                    f += 2.0 * G(p, q, k, k) - G(p, k, k, q);
                if (p == q)
                    e[p] = f;
                else
                    offdiag = std::max(offdiag, std::fabs(f));
            }
        for (int64_t k = 0; k < o; k++)
            E_HF += d.h[k * n + k] + e[k];
        if (offdiag > 1e-6)
            std::printf("warning: the orbitals are not canonical (max |F_pq|, p != q = %.3e)\n", offdiag);
        if (method == "hf") {
            std::cout << std::setprecision(12) << "E(Hartree Fock) = " << E_HF << " hartrees\n";
            for (int64_t p = 0; p < n; p++)
                std::cout << "e[" << p << "] = " << e[p] << "\n";
            return 0;
        }
        // <pq|rs> = (pr|qs), ColMajor n^4 (the layout of TransformTEI(..., true))
        SCESS::CC::TEITensor eri(n, n, n, n);
        for (int64_t s = 0; s < n; s++)
            for (int64_t r = 0; r < n; r++)
                for (int64_t q = 0; q < n; q++)
                    for (int64_t p = 0; p < n; p++)
                        eri(p, q, r, s) = G(p, r, q, s);
        std::stringstream console;
        SCESS::CC::RCCSD_Wfn wfn;
        if (triples) {
            wfn = SCESS::CC::RCCSD_T_FromMO(eri, e, o, E_HF, nullptr, &console);
        } else {
            wfn = SCESS::CC::SolveRCCSD_DIIS(eri, e, o, nullptr, nullptr, &console);
            if (wfn.converged)
                console << "E(Total) = E(Hartree Fock) + E(CCSD) = " << E_HF + wfn.E_CCSD << " hartrees\n\n\n";
        }
        std::cout << std::setprecision(12) << "E(Hartree Fock) = " << E_HF << " hartrees\n" << console.str();
        return wfn.converged ? 0 : 1;
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "scess_cc: %s\n", ex.what());
        return 1;
    }
}
