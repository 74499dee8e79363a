// Input side of the CCSD(T) driver: molecules, BSE-JSON basis sets, AO integrals and closed-shell RHF.
//
// The reference builds these with its Molecule class (src/Quantum/Molecule.cpp:194-294), its BSE-JSON
// basis loader (src/Quantum/Basis/BasisSet.cpp:139-181), libint2 (src/Quantum/Integrals/Integrals.cpp:28-80)
// and RHF with Pulay DIIS on the Fock matrix (src/Quantum/HF/RHF.cpp:34-211). libint2, Eigen and LAPACK
// are not used here: the integrals follow the McMurchie-Davidson scheme (Hermite expansion coefficients
// and Hermite Coulomb integrals from the Boys function, T. Helgaker, P. Jorgensen, J. Olsen, "Molecular
// Electronic-Structure Theory", ch. 9) and the eigenproblems are solved by cyclic Jacobi rotations. Both
// are host code: at the sizes these inputs have (tens of basis functions) they take milliseconds, and the
// n^5 MO transform and everything after it run on the GPU (scess_transform_tei, CCSD, (T)).
#include "../../include/scess_cc.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <dirent.h>

namespace SCESS {

// ------------------------------------------------------------------ molecules
namespace {
constexpr double kAngstrom2Bohr = 1.8897259886;  // src/Quantum/Molecule.h:36

int element_z(const std::string &s) {
    static const char *names[] = {"H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si",
                                  "P", "S", "Cl", "Ar"};
    for (int z = 1; z <= 18; z++)
        if (s == names[z - 1])
            return z;
    throw std::runtime_error("unknown element " + s);
}

Atom make_atom(const std::string &sym, double x, double y, double z) {
    Atom a;
    a.symbol = sym;
    a.Z = element_z(sym);
    a.position[0] = x;
    a.position[1] = y;
    a.position[2] = z;
    return a;
}
}  // namespace

int Molecule::NumElectrons() const {
    int n = 0;
    for (const Atom &a : atoms)
        n += a.Z;
    return n;
}

double Molecule::NuclearRepulsion() const {
    double e = 0.0;
    for (size_t i = 0; i < atoms.size(); i++)
        for (size_t j = 0; j < i; j++) {
            const double dx = atoms[i].position[0] - atoms[j].position[0],
                         dy = atoms[i].position[1] - atoms[j].position[1],
                         dz = atoms[i].position[2] - atoms[j].position[2];
            e += atoms[i].Z * atoms[j].Z / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

// Geometries of the reference's built-in molecules (src/Quantum/Molecule.cpp:194-231, 281-294), in bohr.
Molecule Molecule::GenerateWaterMolecule() {
    const double a = kAngstrom2Bohr;
    Molecule m;
    m.atoms = {make_atom("H", 0.7569520 * a, -0.586349 * a, 0.0), make_atom("O", 0.0, 0.0, 0.0),
               make_atom("H", -0.7569520 * a, -0.586349 * a, 0.0)};
    return m;
}

Molecule Molecule::GenerateHydrogenMolecule() {
    Molecule m;
    m.atoms = {make_atom("H", 0.0, 0.0, 0.0), make_atom("H", 0.7122300374 * kAngstrom2Bohr, 0.0, 0.0)};
    return m;
}

Molecule Molecule::GenerateLithiumHydrideMolecule() {
    Molecule m;
    m.atoms = {make_atom("H", 0.0, 0.0, 0.0), make_atom("Li", 1.60 * kAngstrom2Bohr, 0.0, 0.0)};
    return m;
}

Molecule Molecule::GenerateAmmoniaMolecule() {
    const double a = kAngstrom2Bohr;
    const double N[3] = {-6.6842683864, 1.9896777361, -0.0017590331};
    const double H[3][3] = {{-5.6844990792, 1.9942938061, 0.0072441956},
                            {-7.0127564806, 2.9323453345, 0.0540465122},
                            {-7.0127560539, 1.4847331233, 0.7962183253}};
    Molecule m;
    m.atoms.push_back(make_atom("N", 0.0, 0.0, 0.0));
    for (const auto &h : H)
        m.atoms.push_back(make_atom("H", (h[0] - N[0]) * a, (h[1] - N[1]) * a, (h[2] - N[2]) * a));
    return m;
}

Molecule Molecule::FromName(const std::string &name) {
    if (name == "H2O")
        return GenerateWaterMolecule();
    if (name == "H2")
        return GenerateHydrogenMolecule();
    if (name == "LiH")
        return GenerateLithiumHydrideMolecule();
    if (name == "NH3")
        return GenerateAmmoniaMolecule();
    throw std::runtime_error(name + " is not an available molecule (H2, LiH, NH3, H2O)");
}

// ------------------------------------------------------------------ BSE JSON
namespace {
// A minimal JSON reader: objects, arrays, strings, numbers, true / false / null.
struct JValue {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0.0;
    std::string str;
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;
    const JValue *get(const std::string &k) const {
        for (const auto &kv : obj)
            if (kv.first == k)
                return &kv.second;
        return nullptr;
    }
};

class JParser {
public:
    explicit JParser(const std::string &t) : s_(t) {}
    JValue parse() {
        JValue v = value();
        ws();
        if (i_ != s_.size())
            fail("trailing characters");
        return v;
    }

private:
    const std::string &s_;
    size_t i_ = 0;
    [[noreturn]] void fail(const char *what) {
        throw std::runtime_error(std::string("basis JSON: ") + what + " at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && std::isspace((unsigned char)s_[i_]))
            i_++;
    }
    std::string string_lit() {
        if (s_[i_] != '"')
            fail("expected a string");
        i_++;
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            if (s_[i_] == '\\') {
                if (++i_ >= s_.size())
                    fail("bad escape");
                const char c = s_[i_];
                if (c == 'u') {  // keep the escape as text: basis data never needs the code point
                    out += "\\u";
                } else {
                    out += c == 'n' ? '\n' : c == 't' ? '\t' : c;
                }
            } else {
                out += s_[i_];
            }
            i_++;
        }
        if (i_ >= s_.size())
            fail("unterminated string");
        i_++;
        return out;
    }
    JValue value() {
        ws();
        if (i_ >= s_.size())
            fail("unexpected end");
        JValue v;
        const char c = s_[i_];
        if (c == '{') {
            v.kind = JValue::Obj;
            i_++;
            ws();
            if (s_[i_] == '}') {
                i_++;
                return v;
            }
            for (;;) {
                ws();
                std::string k = string_lit();
                ws();
                if (s_[i_] != ':')
                    fail("expected ':'");
                i_++;
                v.obj.emplace_back(std::move(k), value());
                ws();
                if (s_[i_] == ',') {
                    i_++;
                    continue;
                }
                if (s_[i_] == '}') {
                    i_++;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JValue::Arr;
            i_++;
            ws();
            if (s_[i_] == ']') {
                i_++;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (s_[i_] == ',') {
                    i_++;
                    continue;
                }
                if (s_[i_] == ']') {
                    i_++;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JValue::Str;
            v.str = string_lit();
            return v;
        }
        if (s_.compare(i_, 4, "true") == 0 || s_.compare(i_, 5, "false") == 0 || s_.compare(i_, 4, "null") == 0) {
            v.kind = s_[i_] == 'n' ? JValue::Null : JValue::Bool;
            v.num = s_[i_] == 't';
            i_ += s_[i_] == 'f' ? 5 : 4;
            return v;
        }
        size_t used = 0;
        try {
            v.num = std::stod(s_.substr(i_, 40), &used);
        } catch (...) {
            fail("bad value");
        }
        v.kind = JValue::Num;
        i_ += used;
        return v;
    }
};

double jnum(const JValue &v) {
    if (v.kind == JValue::Num)
        return v.num;
    if (v.kind == JValue::Str)
        return std::stod(v.str);
    throw std::runtime_error("basis JSON: expected a number");
}

std::string lower(std::string s) {
    for (char &c : s)
        c = (char)std::tolower((unsigned char)c);
    return s;
}

bool is_file(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

// "6-31G*" -> "6-31g(d)", "6-31G**" -> "6-31g(d,p)" (the reference accepts both spellings)
std::string canonical_name(const std::string &name) {
    std::string n = lower(name);
    const size_t p2 = n.find("**");
    if (p2 != std::string::npos)
        n = n.substr(0, p2) + "(d,p)" + n.substr(p2 + 2);
    const size_t p1 = n.find('*');
    if (p1 != std::string::npos)
        n = n.substr(0, p1) + "(d)" + n.substr(p1 + 1);
    return n;
}

// <dir>/<name>.json or <dir>/<family>/<name>.json, the name compared case-insensitively
std::string find_basis_file(const std::string &dir, const std::string &name) {
    const std::string want = canonical_name(name) + ".json";
    std::vector<std::string> dirs = {dir};
    if (DIR *d = opendir(dir.c_str())) {
        while (dirent *e = readdir(d)) {
            const std::string n = e->d_name;
            if (n != "." && n != "..") {
                struct stat st;
                const std::string p = dir + "/" + n;
                if (stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode))
                    dirs.push_back(p);
            }
        }
        closedir(d);
    }
    for (const std::string &dd : dirs)
        if (DIR *d = opendir(dd.c_str())) {
            std::string hit;
            while (dirent *e = readdir(d))
                if (lower(e->d_name) == want) {
                    hit = dd + "/" + e->d_name;
                    break;
                }
            closedir(d);
            if (!hit.empty() && is_file(hit))
                return hit;
        }
    return "";
}
}  // namespace

namespace Basis {

BasisSet LoadBasisSet(const std::string &nameOrPath) {
    std::string path = is_file(nameOrPath) ? nameOrPath : "";
    if (path.empty()) {
        const char *dir = std::getenv("SCESS_BASIS_DIR");
        if (!dir)
            throw std::runtime_error(nameOrPath + " is not a file and SCESS_BASIS_DIR is not set");
        path = find_basis_file(dir, nameOrPath);
        if (path.empty())
            throw std::runtime_error(nameOrPath + " is not an available basis set in " + std::string(dir));
    }
    std::ifstream in(path);
    if (!in)
        throw std::runtime_error("cannot open " + path);
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    const JValue root = JParser(text).parse();
    const JValue *elements = root.get("elements");
    if (!elements || elements->kind != JValue::Obj)
        throw std::runtime_error(path + ": no \"elements\" object (BSE JSON expected)");
    BasisSet bs;
    const JValue *nm = root.get("name");
    bs.name = nm && nm->kind == JValue::Str ? nm->str : nameOrPath;
    for (const auto &el : elements->obj) {
        const int Z = std::stoi(el.first);
        const JValue *shells = el.second.get("electron_shells");
        if (!shells)
            continue;
        std::vector<ShellData> &out = bs.elements[Z];
        for (const JValue &sh : shells->arr) {
            const JValue *ft = sh.get("function_type"), *am = sh.get("angular_momentum"), *ex = sh.get("exponents"),
                         *co = sh.get("coefficients");
            if (!am || !ex || !co)
                throw std::runtime_error(path + ": incomplete shell");
            // BasisSet.cpp:157: every shell is spherical unless declared Cartesian
            const bool pure = !(ft && ft->kind == JValue::Str && ft->str == "gto_cartesian");
            std::vector<double> exps;
            for (const JValue &e : ex->arr)
                exps.push_back(jnum(e));
            for (size_t r = 0; r < co->arr.size(); r++) {
                ShellData d;
                // one coefficient row per angular momentum (SP shells), or general contractions of one l
                d.l = (int)jnum(am->arr[std::min(r, am->arr.size() - 1)]);
                if (d.l > 3)
                    throw std::runtime_error(path + ": angular momentum above f is not supported");
                d.pure = pure && d.l >= 2;
                d.exponents = exps;
                for (const JValue &c : co->arr[r].arr)
                    d.coefficients.push_back(jnum(c));
                if (d.coefficients.size() != exps.size())
                    throw std::runtime_error(path + ": coefficient count differs from exponent count");
                out.push_back(std::move(d));
            }
        }
    }
    return bs;
}

}  // namespace Basis

// ------------------------------------------------------------------ integrals
namespace {
constexpr int kLmax = 3;
constexpr int kTmax = 4 * kLmax + 1;  // highest Hermite order of an ERI

// Boys function F_n(T), n = 0..nmax: series for the highest order and downward recursion (T <= 40),
// upward recursion from the error function above (e^-T below 5e-18 there).
void boys(int nmax, double T, double *F) {
    if (T < 1e-14) {
        for (int n = 0; n <= nmax; n++)
            F[n] = 1.0 / (2 * n + 1);
        return;
    }
    if (T > 40.0) {
        F[0] = 0.5 * std::sqrt(M_PI / T) * std::erf(std::sqrt(T));
        const double et = std::exp(-T);
        for (int n = 0; n < nmax; n++)
            F[n + 1] = ((2 * n + 1) * F[n] - et) / (2.0 * T);
        return;
    }
    const double et = std::exp(-T);
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int k = 1; k < 500; k++) {
        term *= 2.0 * T / (2 * nmax + 2 * k + 1);
        sum += term;
        if (term < 1e-17 * sum)
            break;
    }
    F[nmax] = et * sum;
    for (int n = nmax; n > 0; n--)
        F[n - 1] = (2.0 * T * F[n] + et) / (2 * n - 1);
}

// Hermite expansion E[i][j][t] of x_A^i x_B^j for exponents a, b, separation XAB = A - B (one axis)
struct Herm {
    double e[kLmax + 3][kLmax + 3][2 * kLmax + 6];
};
void hermite(int imax, int jmax, double a, double b, double XAB, Herm &H) {
    const double p = a + b, mu = a * b / p;
    const double XPA = -b / p * XAB, XPB = a / p * XAB;
    for (int i = 0; i <= imax; i++)
        for (int j = 0; j <= jmax; j++)
            for (int t = 0; t <= i + j + 2; t++)
                H.e[i][j][t] = 0.0;
    H.e[0][0][0] = std::exp(-mu * XAB * XAB);
    const double inv2p = 0.5 / p;
    for (int i = 0; i <= imax; i++)
        for (int j = 0; j <= jmax; j++) {
            if (i == 0 && j == 0)
                continue;
            for (int t = 0; t <= i + j; t++) {
                double v;
                if (i > 0) {  // raise i
                    const auto &prev = H.e[i - 1][j];
                    v = XPA * prev[t] + (t + 1) * prev[t + 1] + (t > 0 ? inv2p * prev[t - 1] : 0.0);
                } else {  // raise j
                    const auto &prev = H.e[i][j - 1];
                    v = XPB * prev[t] + (t + 1) * prev[t + 1] + (t > 0 ? inv2p * prev[t - 1] : 0.0);
                }
                H.e[i][j][t] = v;
            }
        }
}

// Hermite Coulomb integrals R_{tuv} (order 0) for exponent alpha and separation (X, Y, Z), t+u+v <= nmax,
// by the downward recursion over the auxiliary order n (two levels kept)
constexpr int kR1 = kTmax + 1;
struct HermR {
    double r[kR1 * kR1 * kR1];
    double at(int t, int u, int v) const { return r[(t * kR1 + u) * kR1 + v]; }
};
void hermite_r(int nmax, double alpha, double X, double Y, double Z, HermR &R) {
    static thread_local double lev[2][kR1 * kR1 * kR1];
    auto ix = [](int t, int u, int v) { return (t * kR1 + u) * kR1 + v; };
    double F[kTmax + 2];
    boys(nmax, alpha * (X * X + Y * Y + Z * Z), F);
    double pw[kTmax + 2];
    pw[0] = 1.0;
    for (int n = 1; n <= nmax; n++)
        pw[n] = pw[n - 1] * (-2.0 * alpha);
    double *hi = lev[nmax & 1];
    hi[0] = pw[nmax] * F[nmax];
    for (int n = nmax - 1; n >= 0; n--) {
        const double *up = lev[(n + 1) & 1];
        double *cur = n == 0 ? R.r : lev[n & 1];
        const int top = nmax - n;
        cur[0] = pw[n] * F[n];
        for (int t = 0; t <= top; t++)
            for (int u = 0; u <= top - t; u++)
                for (int v = 0; v <= top - t - u; v++) {
                    if (t + u + v == 0)
                        continue;
                    double val;
                    if (t > 0)
                        val = X * up[ix(t - 1, u, v)] + (t > 1 ? (t - 1) * up[ix(t - 2, u, v)] : 0.0);
                    else if (u > 0)
                        val = Y * up[ix(t, u - 1, v)] + (u > 1 ? (u - 1) * up[ix(t, u - 2, v)] : 0.0);
                    else
                        val = Z * up[ix(t, u, v - 1)] + (v > 1 ? (v - 1) * up[ix(t, u, v - 2)] : 0.0);
                    cur[ix(t, u, v)] = val;
                }
    }
    if (nmax == 0)
        R.r[0] = F[0];
}

struct CShell {  // one contracted Cartesian shell
    int l;
    bool pure;
    double A[3];
    std::vector<double> a, c;  // exponents, coefficients including the x^l primitive normalisation
    int off;                   // first Cartesian function
};

int ncart(int l) { return (l + 1) * (l + 2) / 2; }
void cart_powers(int l, int k, int &lx, int &ly, int &lz) {  // components lx = l..0, ly = l-lx..0
    int idx = 0;
    for (int x = l; x >= 0; x--)
        for (int y = l - x; y >= 0; y--) {
            if (idx++ == k) {
                lx = x;
                ly = y;
                lz = l - x - y;
                return;
            }
        }
}
double dfact(int n) {  // (n)!!, (-1)!! = 1
    double r = 1.0;
    for (int k = n; k > 1; k -= 2)
        r *= k;
    return r;
}

// Cartesian -> real solid harmonic combinations (span of the pure set; scaling is fixed later by the
// overlap diagonal). Rows: pure functions, entries: (cartesian index, weight).
std::vector<std::vector<std::pair<int, double>>> pure_rows(int l) {
    auto idx = [&](int x, int y, int z) {
        for (int k = 0; k < ncart(l); k++) {
            int a, b, c;
            cart_powers(l, k, a, b, c);
            if (a == x && b == y && c == z)
                return k;
        }
        return -1;
    };
    std::vector<std::vector<std::pair<int, double>>> rows;
    if (l == 2) {
        rows = {{{idx(1, 1, 0), 1.0}},
                {{idx(0, 1, 1), 1.0}},
                {{idx(0, 0, 2), 2.0}, {idx(2, 0, 0), -1.0}, {idx(0, 2, 0), -1.0}},
                {{idx(1, 0, 1), 1.0}},
                {{idx(2, 0, 0), 1.0}, {idx(0, 2, 0), -1.0}}};
    } else if (l == 3) {
        rows = {{{idx(0, 0, 3), 2.0}, {idx(2, 0, 1), -3.0}, {idx(0, 2, 1), -3.0}},
                {{idx(1, 0, 2), 4.0}, {idx(3, 0, 0), -1.0}, {idx(1, 2, 0), -1.0}},
                {{idx(0, 1, 2), 4.0}, {idx(2, 1, 0), -1.0}, {idx(0, 3, 0), -1.0}},
                {{idx(2, 0, 1), 1.0}, {idx(0, 2, 1), -1.0}},
                {{idx(1, 1, 1), 1.0}},
                {{idx(3, 0, 0), 1.0}, {idx(1, 2, 0), -3.0}},
                {{idx(2, 1, 0), 3.0}, {idx(0, 3, 0), -1.0}}};
    }
    return rows;
}

std::vector<CShell> molecule_shells(const Molecule &mol, const Basis::BasisSet &bs, int &ncart_total) {
    std::vector<CShell> sh;
    int off = 0;
    for (const Atom &at : mol.atoms) {
        auto it = bs.elements.find(at.Z);
        if (it == bs.elements.end())
            throw std::runtime_error("basis set " + bs.name + " has no entry for " + at.symbol);
        for (const Basis::ShellData &d : it->second) {
            CShell s;
            s.l = d.l;
            s.pure = d.pure;
            for (int k = 0; k < 3; k++)
                s.A[k] = at.position[k];
            s.a = d.exponents;
            for (size_t p = 0; p < d.exponents.size(); p++) {
                const double a = d.exponents[p];
                const double norm = std::pow(2.0 * a / M_PI, 0.75) * std::pow(4.0 * a, 0.5 * d.l) / std::sqrt(dfact(2 * d.l - 1));
                s.c.push_back(d.coefficients[p] * norm);
            }
            s.off = off;
            off += ncart(d.l);
            sh.push_back(std::move(s));
        }
    }
    ncart_total = off;
    return sh;
}
}  // namespace

namespace Integrals {

MolInts EvaluateMolecularIntegrals(const Molecule &mol, const Basis::BasisSet &bs) {
    int nc = 0;
    const std::vector<CShell> sh = molecule_shells(mol, bs, nc);
    if ((double)nc * nc * nc * nc * 8.0 > 4.0e9)
        throw std::runtime_error("EvaluateMolecularIntegrals: more than 4 GB of Cartesian integrals");
    std::vector<double> S((size_t)nc * nc, 0.0), T((size_t)nc * nc, 0.0), V((size_t)nc * nc, 0.0);
    std::vector<double> G((size_t)nc * nc * nc * nc, 0.0);  // (ab|cd), C order over Cartesian functions
    const size_t ns = sh.size();

    // one-electron: overlap, kinetic, nuclear attraction
    for (size_t A = 0; A < ns; A++)
        for (size_t B = 0; B < ns; B++) {
            const CShell &a = sh[A], &b = sh[B];
            for (size_t pa = 0; pa < a.a.size(); pa++)
                for (size_t pb = 0; pb < b.a.size(); pb++) {
                    const double ea = a.a[pa], eb = b.a[pb], p = ea + eb, cc = a.c[pa] * b.c[pb];
                    if (cc == 0.0)
                        continue;
                    Herm Hx, Hy, Hz;
                    hermite(a.l + 1, b.l + 2, ea, eb, a.A[0] - b.A[0], Hx);
                    hermite(a.l + 1, b.l + 2, ea, eb, a.A[1] - b.A[1], Hy);
                    hermite(a.l + 1, b.l + 2, ea, eb, a.A[2] - b.A[2], Hz);
                    const double P[3] = {(ea * a.A[0] + eb * b.A[0]) / p, (ea * a.A[1] + eb * b.A[1]) / p,
                                         (ea * a.A[2] + eb * b.A[2]) / p};
                    const double sq = std::sqrt(M_PI / p);
                    std::vector<HermR> RC(mol.atoms.size());
                    for (size_t k = 0; k < mol.atoms.size(); k++)
                        hermite_r(a.l + b.l, p, P[0] - mol.atoms[k].position[0], P[1] - mol.atoms[k].position[1],
                                  P[2] - mol.atoms[k].position[2], RC[k]);
                    for (int ia = 0; ia < ncart(a.l); ia++)
                        for (int ib = 0; ib < ncart(b.l); ib++) {
                            int ax, ay, az, bx, by, bz;
                            cart_powers(a.l, ia, ax, ay, az);
                            cart_powers(b.l, ib, bx, by, bz);
                            auto s1 = [&](const Herm &H, int i, int j) { return j < 0 ? 0.0 : H.e[i][j][0] * sq; };
                            const double sx = s1(Hx, ax, bx), sy = s1(Hy, ay, by), sz = s1(Hz, az, bz);
                            auto t1 = [&](const Herm &H, int i, int j) {
                                return -0.5 * (j * (j - 1) * s1(H, i, j - 2) - 2.0 * eb * (2 * j + 1) * s1(H, i, j) +
                                               4.0 * eb * eb * s1(H, i, j + 2));
                            };
                            const double kin = t1(Hx, ax, bx) * sy * sz + sx * t1(Hy, ay, by) * sz + sx * sy * t1(Hz, az, bz);
                            double nuc = 0.0;
                            for (size_t k = 0; k < mol.atoms.size(); k++) {
                                double acc = 0.0;
                                for (int t = 0; t <= ax + bx; t++)
                                    for (int u = 0; u <= ay + by; u++)
                                        for (int v = 0; v <= az + bz; v++)
                                            acc += Hx.e[ax][bx][t] * Hy.e[ay][by][u] * Hz.e[az][bz][v] * RC[k].at(t, u, v);
                                nuc -= mol.atoms[k].Z * acc;
                            }
                            const size_t ij = (size_t)(a.off + ia) * nc + (b.off + ib);
                            S[ij] += cc * sx * sy * sz;
                            T[ij] += cc * kin;
                            V[ij] += cc * 2.0 * M_PI / p * nuc;
                        }
                }
        }

    // two-electron: shell quartets with A >= B, C >= D, AB >= CD; all eight images are stored
    auto pair_index = [](size_t x, size_t y) { return x * (x + 1) / 2 + y; };
    for (size_t A = 0; A < ns; A++)
        for (size_t B = 0; B <= A; B++)
            for (size_t C = 0; C < ns; C++)
                for (size_t D = 0; D <= C; D++) {
                    if (pair_index(C, D) > pair_index(A, B))
                        continue;
                    const CShell &a = sh[A], &b = sh[B], &c = sh[C], &d = sh[D];
                    const int na = ncart(a.l), nb = ncart(b.l), ncs = ncart(c.l), nd = ncart(d.l);
                    std::vector<double> q((size_t)na * nb * ncs * nd, 0.0);
                    const int lab = a.l + b.l, lcd = c.l + d.l;
                    for (size_t pa = 0; pa < a.a.size(); pa++)
                        for (size_t pb = 0; pb < b.a.size(); pb++) {
                            const double ea = a.a[pa], eb = b.a[pb], p = ea + eb, cab = a.c[pa] * b.c[pb];
                            if (cab == 0.0)
                                continue;
                            Herm Ex, Ey, Ez;
                            hermite(a.l, b.l, ea, eb, a.A[0] - b.A[0], Ex);
                            hermite(a.l, b.l, ea, eb, a.A[1] - b.A[1], Ey);
                            hermite(a.l, b.l, ea, eb, a.A[2] - b.A[2], Ez);
                            const double P[3] = {(ea * a.A[0] + eb * b.A[0]) / p, (ea * a.A[1] + eb * b.A[1]) / p,
                                                 (ea * a.A[2] + eb * b.A[2]) / p};
                            for (size_t pc = 0; pc < c.a.size(); pc++)
                                for (size_t pd = 0; pd < d.a.size(); pd++) {
                                    const double ec = c.a[pc], ed = d.a[pd], qe = ec + ed, ccd = c.c[pc] * d.c[pd];
                                    if (ccd == 0.0)
                                        continue;
                                    Herm Fx, Fy, Fz;
                                    hermite(c.l, d.l, ec, ed, c.A[0] - d.A[0], Fx);
                                    hermite(c.l, d.l, ec, ed, c.A[1] - d.A[1], Fy);
                                    hermite(c.l, d.l, ec, ed, c.A[2] - d.A[2], Fz);
                                    const double Q[3] = {(ec * c.A[0] + ed * d.A[0]) / qe, (ec * c.A[1] + ed * d.A[1]) / qe,
                                                         (ec * c.A[2] + ed * d.A[2]) / qe};
                                    const double alpha = p * qe / (p + qe);
                                    HermR R;
                                    hermite_r(lab + lcd, alpha, P[0] - Q[0], P[1] - Q[1], P[2] - Q[2], R);
                                    const double pref = cab * ccd * 2.0 * std::pow(M_PI, 2.5) / (p * qe * std::sqrt(p + qe));
                                    for (int ic = 0; ic < ncs; ic++)
                                        for (int id = 0; id < nd; id++) {
                                            int cx, cy, cz, dx, dy, dz;
                                            cart_powers(c.l, ic, cx, cy, cz);
                                            cart_powers(d.l, id, dx, dy, dz);
                                            // inner sum over the (cd) Hermite indices for every (t, u, v) of (ab)
                                            double inner[2 * kLmax + 1][2 * kLmax + 1][2 * kLmax + 1];
                                            for (int t = 0; t <= lab; t++)
                                                for (int u = 0; u <= lab - t; u++)
                                                    for (int v = 0; v <= lab - t - u; v++) {
                                                        double s = 0.0;
                                                        for (int tt = 0; tt <= cx + dx; tt++)
                                                            for (int uu = 0; uu <= cy + dy; uu++)
                                                                for (int vv = 0; vv <= cz + dz; vv++) {
                                                                    const double e = Fx.e[cx][dx][tt] * Fy.e[cy][dy][uu] * Fz.e[cz][dz][vv];
                                                                    s += ((tt + uu + vv) & 1 ? -e : e) * R.at(t + tt, u + uu, v + vv);
                                                                }
                                                        inner[t][u][v] = s;
                                                    }
                                            for (int ia = 0; ia < na; ia++)
                                                for (int ib = 0; ib < nb; ib++) {
                                                    int ax, ay, az, bx, by, bz;
                                                    cart_powers(a.l, ia, ax, ay, az);
                                                    cart_powers(b.l, ib, bx, by, bz);
                                                    double s = 0.0;
                                                    for (int t = 0; t <= ax + bx; t++)
                                                        for (int u = 0; u <= ay + by; u++)
                                                            for (int v = 0; v <= az + bz; v++)
                                                                s += Ex.e[ax][bx][t] * Ey.e[ay][by][u] * Ez.e[az][bz][v] * inner[t][u][v];
                                                    q[(((size_t)ia * nb + ib) * ncs + ic) * nd + id] += pref * s;
                                                }
                                        }
                                }
                        }
                    for (int ia = 0; ia < na; ia++)
                        for (int ib = 0; ib < nb; ib++)
                            for (int ic = 0; ic < ncs; ic++)
                                for (int id = 0; id < nd; id++) {
                                    const double x = q[(((size_t)ia * nb + ib) * ncs + ic) * nd + id];
                                    const size_t i = a.off + ia, j = b.off + ib, k = c.off + ic, l = d.off + id;
                                    auto put = [&](size_t p0, size_t p1, size_t p2, size_t p3) {
                                        G[((p0 * nc + p1) * nc + p2) * nc + p3] = x;
                                    };
                                    put(i, j, k, l), put(j, i, k, l), put(i, j, l, k), put(j, i, l, k);
                                    put(k, l, i, j), put(l, k, i, j), put(k, l, j, i), put(l, k, j, i);
                                }
                }

    // Cartesian -> basis functions (pure shells to real solid harmonics), then unit norms
    std::vector<std::vector<std::pair<int, double>>> cols;  // basis function -> (cartesian index, weight)
    for (const CShell &s : sh) {
        if (s.pure) {
            for (auto &row : pure_rows(s.l)) {
                std::vector<std::pair<int, double>> c;
                for (auto &e : row)
                    c.emplace_back(s.off + e.first, e.second);
                cols.push_back(c);
            }
        } else {
            for (int k = 0; k < ncart(s.l); k++)
                cols.push_back({{s.off + k, 1.0}});
        }
    }
    const int n = (int)cols.size();
    std::vector<double> M((size_t)nc * n, 0.0);  // nc x n
    for (int j = 0; j < n; j++)
        for (auto &e : cols[j])
            M[(size_t)e.first * n + j] = e.second;
    auto one_e = [&](const std::vector<double> &X) {  // M^T X M
        std::vector<double> t((size_t)nc * n, 0.0), r((size_t)n * n, 0.0);
        for (int p = 0; p < nc; p++)
            for (int j = 0; j < n; j++) {
                double s = 0.0;
                for (int q2 = 0; q2 < nc; q2++)
                    s += X[(size_t)p * nc + q2] * M[(size_t)q2 * n + j];
                t[(size_t)p * n + j] = s;
            }
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double s = 0.0;
                for (int p = 0; p < nc; p++)
                    s += M[(size_t)p * n + i] * t[(size_t)p * n + j];
                r[(size_t)i * n + j] = s;
            }
        return r;
    };
    std::vector<double> Sb = one_e(S), Tb = one_e(T), Vb = one_e(V);
    std::vector<double> scale(n);
    for (int i = 0; i < n; i++)
        scale[i] = 1.0 / std::sqrt(Sb[(size_t)i * n + i]);
    // four quarter transformations of G over its last index, rotating the indices each time
    std::vector<double> cur = G, nxt;
    int dims[4] = {nc, nc, nc, nc};
    for (int pass = 0; pass < 4; pass++) {
        const size_t lead = (size_t)dims[1] * dims[2] * dims[3];
        const int d0 = dims[0];
        // cur[a][b][c][d] (a over d0) -> nxt[b][c][d][j] = sum_a cur[a][b][c][d] M[a][j] * scale[j]
        nxt.assign(lead * n, 0.0);
        for (int a = 0; a < d0; a++)
            for (int j = 0; j < n; j++) {
                const double w = M[(size_t)a * n + j] * scale[j];
                if (w == 0.0)
                    continue;
                const double *src = cur.data() + (size_t)a * lead;
                for (size_t r = 0; r < lead; r++)
                    nxt[r * n + j] += w * src[r];
            }
        cur.swap(nxt);
        dims[0] = dims[1];
        dims[1] = dims[2];
        dims[2] = dims[3];
        dims[3] = n;
    }
    MolInts mi;
    mi.S = Tensor<2>(n, n);
    mi.T = Tensor<2>(n, n);
    mi.V = Tensor<2>(n, n);
    mi.eri = TEITensor(n, n, n, n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            const double f = scale[i] * scale[j];
            mi.S(i, j) = Sb[(size_t)i * n + j] * f;
            mi.T(i, j) = Tb[(size_t)i * n + j] * f;
            mi.V(i, j) = Vb[(size_t)i * n + j] * f;
        }
    // cur is C order over (i, j, k, l); the ColMajor tensor's (i, j, k, l) has the same value
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++)
                for (int l = 0; l < n; l++)
                    mi.eri(i, j, k, l) = cur[(((size_t)i * n + j) * n + k) * n + l];
    return mi;
}

}  // namespace Integrals

// ------------------------------------------------------------------ RHF
namespace {
// Symmetric eigenproblem A = V diag(w) V^T by cyclic Jacobi rotations (A row-major n x n, destroyed);
// eigenvalues ascending, V row-major with eigenvectors in columns.
void jacobi_eigen(std::vector<double> A, int n, std::vector<double> &w, std::vector<double> &V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        V[(size_t)i * n + i] = 1.0;
    auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += a(i, j) * a(i, j);
                if (i != j)
                    off += a(i, j) * a(i, j);
            }
        if (off <= 1e-30 * std::max(tot, 1e-300))
            break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = a(p, q);
                if (std::fabs(apq) < 1e-300)
                    continue;
                const double theta = (a(q, q) - a(p, p)) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {  // columns p, q
                    const double akp = a(k, p), akq = a(k, q);
                    a(k, p) = c * akp - s * akq;
                    a(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {  // rows p, q
                    const double apk = a(p, k), aqk = a(q, k);
                    a(p, k) = c * apk - s * aqk;
                    a(q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; i++)
        ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
    std::vector<double> Vs((size_t)n * n);
    w.resize(n);
    for (int j = 0; j < n; j++) {
        w[j] = a(ord[j], ord[j]);
        for (int i = 0; i < n; i++)
            Vs[(size_t)i * n + j] = V[(size_t)i * n + ord[j]];
    }
    V.swap(Vs);
}

std::vector<double> matmul(const std::vector<double> &A, const std::vector<double> &B, int n, bool ta = false,
                           bool tb = false) {
    std::vector<double> C((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) {
            const double aik = ta ? A[(size_t)k * n + i] : A[(size_t)i * n + k];
            if (aik == 0.0)
                continue;
            for (int j = 0; j < n; j++)
                C[(size_t)i * n + j] += aik * (tb ? B[(size_t)j * n + k] : B[(size_t)k * n + j]);
        }
    return C;
}

// Solves the symmetric system M x = r (m x m) by Gaussian elimination with partial pivoting; an empty
// result when M is singular
std::vector<double> solve(std::vector<double> M, std::vector<double> r, int m) {
    for (int c = 0; c < m; c++) {
        int piv = c;
        for (int i = c + 1; i < m; i++)
            if (std::fabs(M[(size_t)i * m + c]) > std::fabs(M[(size_t)piv * m + c]))
                piv = i;
        for (int j = 0; j < m; j++)
            std::swap(M[(size_t)c * m + j], M[(size_t)piv * m + j]);
        std::swap(r[c], r[piv]);
        const double d = M[(size_t)c * m + c];
        if (std::fabs(d) < 1e-300)
            return {};
        for (int i = c + 1; i < m; i++) {
            const double f = M[(size_t)i * m + c] / d;
            for (int j = c; j < m; j++)
                M[(size_t)i * m + j] -= f * M[(size_t)c * m + j];
            r[i] -= f * r[c];
        }
    }
    std::vector<double> x(m);
    for (int i = m - 1; i >= 0; i--) {
        double s = r[i];
        for (int j = i + 1; j < m; j++)
            s -= M[(size_t)i * m + j] * x[j];
        x[i] = s / M[(size_t)i * m + i];
    }
    return x;
}
}  // namespace

namespace HF {

HF_Wfn RHF(const Molecule &molecule, const Basis::BasisSet &basis, double convergeLimit, size_t scfMaxIterations,
           Integrals::MolInts *mintsOut, const Thread *thread, std::stringstream *console) {
    const int nel = molecule.NumElectrons();
    if (nel % 2)
        throw std::runtime_error("RHF needs an even number of electrons");
    const int nocc = nel / 2;
    Integrals::MolInts mi = Integrals::EvaluateMolecularIntegrals(molecule, basis);
    const int n = (int)mi.S.dimension(0);
    if (nocc > n)
        throw std::runtime_error("RHF: fewer basis functions than occupied orbitals");
    const double Vnn = molecule.NuclearRepulsion();
    std::vector<double> S((size_t)n * n), H((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            S[(size_t)i * n + j] = mi.S(i, j);
            H[(size_t)i * n + j] = mi.T(i, j) + mi.V(i, j);
        }
    // X = S^-1/2 (symmetric orthogonalisation)
    std::vector<double> sw, sv;
    jacobi_eigen(S, n, sw, sv);
    if (sw[0] < 1e-10)
        throw std::runtime_error("RHF: the basis is linearly dependent");
    std::vector<double> X((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            for (int k = 0; k < n; k++)
                s += sv[(size_t)i * n + k] * sv[(size_t)j * n + k] / std::sqrt(sw[k]);
            X[(size_t)i * n + j] = s;
        }
    auto diagonalise = [&](const std::vector<double> &F, std::vector<double> &eps, std::vector<double> &C) {
        const std::vector<double> Fp = matmul(matmul(X, F, n, true), X, n);  // X^T F X
        std::vector<double> Cp;
        jacobi_eigen(Fp, n, eps, Cp);
        C = matmul(X, Cp, n);
    };
    auto density = [&](const std::vector<double> &C) {  // D = 2 C_occ C_occ^T
        std::vector<double> D((size_t)n * n, 0.0);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double s = 0.0;
                for (int k = 0; k < nocc; k++)
                    s += C[(size_t)i * n + k] * C[(size_t)j * n + k];
                D[(size_t)i * n + j] = 2.0 * s;
            }
        return D;
    };
    auto fock = [&](const std::vector<double> &D) {  // F = H + J - K/2
        std::vector<double> F = H;
        for (int m = 0; m < n; m++)
            for (int nu = 0; nu < n; nu++) {
                double g = 0.0;
                for (int l = 0; l < n; l++)
                    for (int s = 0; s < n; s++)
                        g += D[(size_t)l * n + s] * (mi.eri(m, nu, l, s) - 0.5 * mi.eri(m, l, nu, s));
                F[(size_t)m * n + nu] += g;
            }
        return F;
    };
    auto energy = [&](const std::vector<double> &D, const std::vector<double> &F) {
        double e = 0.0;
        for (size_t k = 0; k < D.size(); k++)
            e += 0.5 * D[k] * (H[k] + F[k]);
        return e + Vnn;
    };

    std::vector<double> eps, C;
    diagonalise(H, eps, C);  // core-Hamiltonian guess
    std::vector<double> D = density(C);
    double E = 0.0, Elast = 0.0;
    bool converged = false;
    std::vector<std::vector<double>> fs, es;  // DIIS: Fock matrices and errors X^T (FDS - SDF) X
    const double commutatorLimit = std::sqrt(convergeLimit);
    for (size_t it = 0; it < scfMaxIterations; it++) {
        if (thread && thread->ShouldStop())
            return HF_Wfn{};
        std::vector<double> F = fock(D);
        E = energy(D, F);
        const std::vector<double> FDS = matmul(matmul(F, D, n), S, n), SDF = matmul(matmul(S, D, n), F, n);
        std::vector<double> err((size_t)n * n);
        for (size_t k = 0; k < err.size(); k++)
            err[k] = FDS[k] - SDF[k];
        err = matmul(matmul(X, err, n, true), X, n);
        double rms = 0.0;
        for (double x : err)
            rms += x * x;
        rms = std::sqrt(rms / (double)err.size());
        if (console)
            *console << "SCF(" << it << "): E = " << E << " rms([F,D]) = " << rms << "\n";
        if (it > 1 && rms < commutatorLimit && std::fabs(E - Elast) < convergeLimit) {
            converged = true;
            break;
        }
        Elast = E;
        fs.push_back(F);
        es.push_back(err);
        if (fs.size() > 6) {
            fs.erase(fs.begin());
            es.erase(es.begin());
        }
        if (fs.size() >= 2) {  // Pulay extrapolation of F (the B block scaled by its largest entry)
            const int m = (int)fs.size() + 1;
            std::vector<double> B((size_t)m * m, 0.0), r(m, 0.0);
            double bmax = 0.0;
            for (int i = 0; i < m - 1; i++)
                for (int j = 0; j < m - 1; j++) {
                    double s = 0.0;
                    for (size_t k = 0; k < es[i].size(); k++)
                        s += es[i][k] * es[j][k];
                    B[(size_t)i * m + j] = s;
                    bmax = std::max(bmax, std::fabs(s));
                }
            if (bmax >= 1e-28) {  // else the commutator vanishes (e.g. by symmetry): nothing to extrapolate
                for (int i = 0; i < m - 1; i++)
                    for (int j = 0; j < m - 1; j++)
                        B[(size_t)i * m + j] /= bmax;
                for (int i = 0; i < m - 1; i++)
                    B[(size_t)i * m + m - 1] = B[(size_t)(m - 1) * m + i] = -1.0;
                r[m - 1] = -1.0;
                // a singular system (two equal error vectors, say) leaves this iteration's Fock matrix as it is
                const std::vector<double> cdiis = solve(B, r, m);
                if (!cdiis.empty()) {
                    std::fill(F.begin(), F.end(), 0.0);
                    for (int i = 0; i < m - 1; i++)
                        for (size_t k = 0; k < F.size(); k++)
                            F[k] += cdiis[i] * fs[i][k];
                }
            }
        }
        diagonalise(F, eps, C);
        D = density(C);
    }
    if (!converged)
        throw std::runtime_error("RHF did not converge in " + std::to_string(scfMaxIterations) + " iterations");
    // final orbitals from the converged Fock matrix
    diagonalise(fock(D), eps, C);
    HF_Wfn w;
    w.E_HF = E;
    w.nao = n;
    w.nmo = n;
    w.C_Alpha.assign((size_t)n * n, 0.0);
    for (int p = 0; p < n; p++)
        for (int mu = 0; mu < n; mu++)
            w.C_Alpha[(size_t)p * n + mu] = C[(size_t)mu * n + p];  // ColMajor nao x nmo
    w.Ep_Alpha = eps;
    w.Occ_Alpha.assign(n, 0);
    for (int k = 0; k < nocc; k++)
        w.Occ_Alpha[k] = 1;
    if (mintsOut)
        *mintsOut = std::move(mi);
    return w;
}

}  // namespace HF
}  // namespace SCESS
