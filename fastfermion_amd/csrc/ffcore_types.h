// ffcore_types.h -- host-side operator types of the drop-in module.
//
// These restate the value semantics of the reference types that the
// `propagate` path consumes and produces (src/pauli/algebra.h,
// src/pauli/gates.h, src/majorana_algebra.h, src/majorana_gates.h,
// src/qubitproductstate.h, src/fockstate.h).  Keys use 4 words per mask:
// Pauli strings on up to 256 qubits, Majorana strings on up to 256 Majorana
// operators.  Propagation itself is not done here: it goes through the C ABI
// to the device engine.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace ffh {

constexpr int PW = 4;
constexpr int MW = 4;
constexpr int MAX_QUBITS = 64 * PW;
constexpr int MAX_MAJORANAS = 64 * MW;
using cplx = std::complex<double>;

inline int popc(uint64_t v) { return __builtin_popcountll(v); }

inline uint64_t mix64(uint64_t h) {
    h ^= h >> 30;
    h *= 0xbf58476d1ce4e5b9ull;
    h ^= h >> 27;
    h *= 0x94d049bb133111ebull;
    h ^= h >> 31;
    return h;
}

// i^m for any integer m (src/common.h:57-66)
inline cplx ipow(int m) {
    const int r = ((m % 4) + 4) % 4;
    switch (r) {
        case 0: return cplx(1, 0);
        case 1: return cplx(0, 1);
        case 2: return cplx(-1, 0);
        default: return cplx(0, -1);
    }
}

// The reference evaluates cos and sin of one angle side by side; GCC -O2
// makes that one glibc sincos() call, so the host does the same.
inline void trig(double t, double& c, double& s) { ::sincos(t, &s, &c); }

[[noreturn]] inline void fail(const std::string& msg) { throw std::runtime_error(msg); }

// ---------------------------------------------------------------------------
// Pauli strings (src/pauli/algebra.h:30-223): xory / yorz masks.
// ---------------------------------------------------------------------------
struct PauliString {
    uint64_t x[PW] = {0, 0, 0, 0};
    uint64_t z[PW] = {0, 0, 0, 0};

    bool operator==(const PauliString& o) const {
        return std::memcmp(x, o.x, sizeof x) == 0 && std::memcmp(z, o.z, sizeof z) == 0;
    }
    bool is_identity() const {
        for (int w = 0; w < PW; ++w)
            if (x[w] || z[w]) return false;
        return true;
    }
    uint64_t hash() const {
        uint64_t h = 0x51ed270b27a3f4c5ull;
        for (int w = 0; w < PW; ++w) h = mix64(h ^ x[w]) + z[w];
        return mix64(h);
    }
    bool xbit(int q) const { return (x[q >> 6] >> (q & 63)) & 1ull; }
    bool zbit(int q) const { return (z[q >> 6] >> (q & 63)) & 1ull; }
    void set(int q, char op) {
        const uint64_t m = 1ull << (q & 63);
        if (op == 'X' || op == 'Y') x[q >> 6] |= m;
        if (op == 'Y' || op == 'Z') z[q >> 6] |= m;
    }
    int degree_x() const {
        int d = 0;
        for (int w = 0; w < PW; ++w) d += popc(x[w] & ~z[w]);
        return d;
    }
    int degree_y() const {
        int d = 0;
        for (int w = 0; w < PW; ++w) d += popc(x[w] & z[w]);
        return d;
    }
    int degree_z() const {
        int d = 0;
        for (int w = 0; w < PW; ++w) d += popc(~x[w] & z[w]);
        return d;
    }
    int degree() const {
        int d = 0;
        for (int w = 0; w < PW; ++w) d += popc(x[w] | z[w]);
        return d;
    }
    int degree(char v) const {
        if (v == 'X' || v == 'x') return degree_x();
        if (v == 'Y' || v == 'y') return degree_y();
        if (v == 'Z' || v == 'z') return degree_z();
        if (v == '\0') return degree();
        fail("Invalid argument: should be either 'X', 'Y', or 'Z' or empty char");
    }
    // index of the highest non-identity qubit + 1 (src/pauli/algebra.h:180)
    int extent() const {
        for (int w = PW - 1; w >= 0; --w) {
            const uint64_t s = x[w] | z[w];
            if (s) return 64 * w + 64 - __builtin_clzll(s);
        }
        return 0;
    }
    std::vector<std::pair<int, char>> indices() const {
        std::vector<std::pair<int, char>> r;
        for (int q = 0; q < MAX_QUBITS; ++q) {
            const bool a = xbit(q), b = zbit(q);
            if (a && b)
                r.emplace_back(q, 'Y');
            else if (a)
                r.emplace_back(q, 'X');
            else if (b)
                r.emplace_back(q, 'Z');
        }
        return r;
    }
    std::vector<int> support() const {
        std::vector<int> r;
        for (auto& p : indices()) r.push_back(p.first);
        return r;
    }
    std::string to_string(int n = 0) const {
        const int len = std::max(extent(), n);
        std::string s(len, 'I');
        for (auto& p : indices()) s[p.first] = p.second;
        return s;
    }
    std::string compact() const {
        if (is_identity()) return "I";
        std::string s;
        for (auto& p : indices()) {
            if (!s.empty()) s += " ";
            s += p.second + std::to_string(p.first);
        }
        return s;
    }
    // PauliString::commutes (src/pauli/algebra.h:209-211)
    bool commutes(const PauliString& b) const {
        int p1 = 0, p2 = 0;
        for (int w = 0; w < PW; ++w) {
            p1 += popc(x[w] & b.z[w]);
            p2 += popc(z[w] & b.x[w]);
        }
        return (p1 & 1) == (p2 & 1);
    }
    // highest bit index + 1 over both masks, for choosing the device word count
    int bits_used() const { return extent(); }
};

struct PSHash {
    size_t operator()(const PauliString& a) const { return (size_t)a.hash(); }
};

// Product a*b = i^jpow * r (src/pauli/algebra.h:227-233, 577-582)
inline PauliString pauli_mul(const PauliString& a, const PauliString& b, int& jpow) {
    PauliString r;
    int j = 0;
    for (int w = 0; w < PW; ++w) j += popc(a.x[w] & a.z[w]) + popc(b.x[w] & b.z[w]) + 2 * popc(a.z[w] & b.x[w]);
    for (int w = 0; w < PW; ++w) {
        r.x[w] = a.x[w] ^ b.x[w];
        r.z[w] = a.z[w] ^ b.z[w];
    }
    for (int w = 0; w < PW; ++w) j -= popc(r.x[w] & r.z[w]);
    jpow = j;
    return r;
}

// "XXIYZ" or "X0 Y1 Z4" (src/pauli/algebra.h:74-98, src/parse.h:18-47)
inline std::vector<std::pair<int, char>> parse_pauli_indexed(const std::string& ps) {
    std::vector<std::pair<int, char>> ind;
    if (ps == "I" || ps == "1") return ind;
    const int n = (int)ps.size();
    for (int i = 0; i < n;) {
        if (i > 0 && (ps[i] == ' ' || ps[i] == '*')) {
            i++;
            continue;
        }
        if (ps[i] == 'X' || ps[i] == 'Y' || ps[i] == 'Z') {
            int j = i + 1;
            while (j < n && std::isdigit((unsigned char)ps[j])) j++;
            if (j == i + 1) fail("Invalid Pauli string");
            ind.emplace_back(std::stoi(ps.substr(i + 1, j - i - 1)), ps[i]);
            i = j;
        } else {
            fail("Invalid Pauli string");
        }
    }
    return ind;
}

inline PauliString pauli_from_indices(const std::vector<std::pair<int, char>>& s) {
    PauliString p;
    for (size_t i = 0; i < s.size(); ++i) {
        for (size_t j = 0; j < s.size(); ++j)
            if (j != i && s[j].first == s[i].first)
                fail("Invalid PauliString (qubit index " + std::to_string(s[i].first) + " appears multiple times)");
        const int q = s[i].first;
        const char op = s[i].second;
        if (q < 0 || q >= MAX_QUBITS)
            fail("Invalid PauliString (qubit index " + std::to_string(q) + " invalid, should be between 0 and " +
                 std::to_string(MAX_QUBITS - 1) + ")");
        if (op != 'I' && op != 'X' && op != 'Y' && op != 'Z')
            fail(std::string("Invalid PauliString (symbol '") + op + "' not recognized)");
        p.set(q, op);
    }
    return p;
}

inline PauliString pauli_from_string(const std::string& s) {
    bool dense = true;
    for (char c : s)
        if (!(c == 'X' || c == 'Y' || c == 'Z' || c == 'I')) {
            dense = false;
            break;
        }
    if (dense) {
        if ((int)s.size() > MAX_QUBITS) fail("Invalid PauliString (too many qubits)");
        PauliString p;
        for (int i = 0; i < (int)s.size(); ++i) p.set(i, s[i]);
        return p;
    }
    return pauli_from_indices(parse_pauli_indexed(s));
}

// ---------------------------------------------------------------------------
// Majorana strings (src/majorana_algebra.h:30-106)
// ---------------------------------------------------------------------------
struct MajoranaString {
    uint64_t m[MW] = {0, 0, 0, 0};
    bool operator==(const MajoranaString& o) const { return std::memcmp(m, o.m, sizeof m) == 0; }
    bool is_identity() const { return !(m[0] | m[1] | m[2] | m[3]); }
    uint64_t hash() const {
        uint64_t h = 0x2545f4914f6cdd1dull;
        for (int w = 0; w < MW; ++w) h = mix64(h ^ m[w]);
        return h;
    }
    bool bit(int i) const { return (m[i >> 6] >> (i & 63)) & 1ull; }
    int degree() const {
        int d = 0;
        for (int w = 0; w < MW; ++w) d += popc(m[w]);
        return d;
    }
    // deg mod 4 in {0, 1} (src/majorana_algebra.h:68-71)
    bool is_hermitian() const {
        const int d = degree() % 4;
        return d == 0 || d == 1;
    }
    int extent() const {
        for (int w = MW - 1; w >= 0; --w)
            if (m[w]) return 64 * w + 64 - __builtin_clzll(m[w]);
        return 0;
    }
    std::vector<int> support() const {
        std::vector<int> r;
        for (int i = 0; i < MAX_MAJORANAS; ++i)
            if (bit(i)) r.push_back(i);
        return r;
    }
    std::string compact() const {
        std::string s;
        for (int i : support()) {
            if (!s.empty()) s += " ";
            s += "m" + std::to_string(i);
        }
        return s.empty() ? "I" : s;
    }
    // (|a||b| - |a&b|) even (src/majorana_algebra.h:79-84)
    bool commutes(const MajoranaString& b) const {
        int nab = 0;
        for (int w = 0; w < MW; ++w) nab += popc(m[w] & b.m[w]);
        return ((degree() * b.degree() - nab) % 2) == 0;
    }
};

struct MSHash {
    size_t operator()(const MajoranaString& a) const { return (size_t)a.hash(); }
};

inline MajoranaString majorana_from_support(const std::vector<int>& supp) {
    MajoranaString s;
    for (size_t i = 0; i < supp.size(); ++i) {
        if (i > 0 && supp[i] <= supp[i - 1]) fail("Invalid MajoranaString (not normal ordered)");
        if (supp[i] < 0 || supp[i] >= MAX_MAJORANAS)
            fail("Invalid MajoranaString (mode index " + std::to_string(supp[i]) + " invalid, should be between 0 and " +
                 std::to_string(MAX_MAJORANAS - 1) + ")");
        s.m[supp[i] >> 6] |= 1ull << (supp[i] & 63);
    }
    return s;
}

// parity of #{(i in x, j in y): i < j} (src/bits_utils.h:14-38)
inline int merge_parity(const MajoranaString& x, const MajoranaString& y) {
    int tot = 0;
    int below = 0;  // bits of x strictly below the current position
    for (int i = 0; i < MAX_MAJORANAS; ++i) {
        if (y.bit(i)) tot += below;
        if (x.bit(i)) ++below;
    }
    return tot & 1;
}

// x*y = sign * (x ^ y) (src/majorana_algebra.h:441-444)
inline MajoranaString majorana_mul(const MajoranaString& x, const MajoranaString& y, int& sign) {
    MajoranaString r;
    for (int w = 0; w < MW; ++w) r.m[w] = x.m[w] ^ y.m[w];
    sign = merge_parity(y, x) ? -1 : 1;
    return r;
}

// ---------------------------------------------------------------------------
// Insertion-ordered polynomial container with erase-by-swap (the iteration
// semantics of the reference's map, src/hashmap/unordered_dense.h).
// ---------------------------------------------------------------------------
template <class K, class H>
struct Terms {
    std::vector<K> keys;
    std::vector<cplx> vals;
    std::unordered_map<K, uint32_t, H> idx;

    size_t size() const { return keys.size(); }
    void reserve(size_t n) {
        keys.reserve(n);
        vals.reserve(n);
        idx.reserve(n);
    }
    cplx& at(const K& k) {
        auto it = idx.find(k);
        if (it != idx.end()) return vals[it->second];
        idx.emplace(k, (uint32_t)keys.size());
        keys.push_back(k);
        vals.push_back(cplx(0, 0));
        return vals.back();
    }
    const cplx* find(const K& k) const {
        auto it = idx.find(k);
        return it == idx.end() ? nullptr : &vals[it->second];
    }
    void erase(const K& k) {
        auto it = idx.find(k);
        if (it == idx.end()) return;
        const uint32_t i = it->second;
        idx.erase(it);
        const uint32_t last = (uint32_t)keys.size() - 1;
        if (i != last) {
            keys[i] = keys[last];
            vals[i] = vals[last];
            idx[keys[i]] = i;
        }
        keys.pop_back();
        vals.pop_back();
    }
    // add and drop an exact zero (the reference's += then inplace_compress)
    void add_compress(const K& k, cplx v) {
        cplx& r = at(k);
        r += v;
        if (r == cplx(0, 0)) erase(k);
    }
    void clear() {
        keys.clear();
        vals.clear();
        idx.clear();
    }
};

using PauliTerms = Terms<PauliString, PSHash>;
using MajoranaTerms = Terms<MajoranaString, MSHash>;

struct PauliPolynomial {
    PauliTerms t;
};

struct MajoranaPolynomial {
    MajoranaTerms t;
};

// ---------------------------------------------------------------------------
// Gates (src/pauli/gates.h, src/majorana_gates.h)
// ---------------------------------------------------------------------------
enum CliffKind { kH = 1, kS = 2, kCNOT = 3, kSWAP = 4, kCZ = 5 };

inline void set_bit(uint64_t* v, int q, bool b) {
    const uint64_t m = 1ull << (q & 63);
    v[q >> 6] = b ? (v[q >> 6] | m) : (v[q >> 6] & ~m);
}

// H_impl .. CZ_impl (src/pauli/gates.h:15-69)
inline void clifford_apply(int kind, int i, int j, PauliString& a, cplx& c) {
    const bool xi = a.xbit(i), zi = a.zbit(i);
    switch (kind) {
        case kH:
            set_bit(a.x, i, zi);
            set_bit(a.z, i, xi);
            if (xi && zi) c *= -1;
            break;
        case kS:
            set_bit(a.z, i, xi ^ zi);
            if (xi && !zi) c *= -1;
            break;
        case kCNOT: {
            const bool xj = a.xbit(j), zj = a.zbit(j);
            set_bit(a.x, j, xi ^ xj);
            set_bit(a.z, i, zi ^ zj);
            if (xi && zj && !(zi ^ xj)) c *= -1;
            break;
        }
        case kSWAP: {
            const bool xj = a.xbit(j), zj = a.zbit(j);
            set_bit(a.x, i, xj);
            set_bit(a.x, j, xi);
            set_bit(a.z, i, zj);
            set_bit(a.z, j, zi);
            break;
        }
        case kCZ: {
            const bool xj = a.xbit(j), zj = a.zbit(j);
            if (xi && xj && (zi ^ zj)) c *= -1;
            set_bit(a.z, i, zi ^ xj);
            set_bit(a.z, j, zj ^ xi);
            break;
        }
        default:
            fail("unknown Clifford gate");
    }
}

struct GateH {
    int i;
};
struct GateS {
    int i;
};
struct GateCNOT {
    int i, j;
};
struct GateSWAP {
    int i, j;
};
struct GateCZ {
    int i, j;
};

struct RotGate {
    PauliString ps;
    double theta;
};

struct MRotGate {
    MajoranaString ms;
    double theta;
    bool r;
};

// ---------------------------------------------------------------------------
// Jordan-Wigner bridge between Majorana and Pauli strings (src/transforms.h:125-160,
// 254-298), with mode j <-> qubit j and
//   m_{2j} = Z_0 ... Z_{j-1} X_j,   m_{2j+1} = Z_0 ... Z_{j-1} Y_j.
// ---------------------------------------------------------------------------

// Image of a normal-ordered Majorana string: the product of the images of its
// factors in ascending index order; returns the Pauli string, phase in `phase`.
inline PauliString jw_string(const MajoranaString& a, cplx& phase) {
    PauliString res;
    int jpow = 0;
    for (int i : a.support()) {
        const int j = i >> 1;
        if (j >= MAX_QUBITS) fail("Majorana index out of the Pauli range");
        PauliString f;
        for (int q = 0; q < j; ++q) f.set(q, 'Z');
        f.set(j, (i & 1) ? 'Y' : 'X');
        int jj = 0;
        res = pauli_mul(res, f, jj);
        jpow += jj;
    }
    phase = ipow(jpow);
    return res;
}

// Inverse map.  From the definitions: Z_j = -i m_{2j} m_{2j+1}, so
// Z_0 ... Z_{j-1} = (-i)^j m_0 m_1 ... m_{2j-1}, X_j = (-i)^j m_0 ... m_{2j-1} m_{2j}
// and Y_j = (-i)^j m_0 ... m_{2j-1} m_{2j+1}; the images of different qubits
// commute, so the string is their product in any order.
inline MajoranaString majorana_of_pauli(const PauliString& p, cplx& phase) {
    MajoranaString res;
    int jpow = 0;
    int sign = 1;
    for (const auto& [j, op] : p.indices()) {
        if (2 * j + 1 >= MAX_MAJORANAS) fail("qubit index out of the Majorana range");
        MajoranaString f;
        if (op == 'Z') {
            f.m[(2 * j) >> 6] |= 1ull << ((2 * j) & 63);
            f.m[(2 * j + 1) >> 6] |= 1ull << ((2 * j + 1) & 63);
            jpow -= 1;
        } else {
            for (int i = 0; i < 2 * j; ++i) f.m[i >> 6] |= 1ull << (i & 63);
            const int last = op == 'X' ? 2 * j : 2 * j + 1;
            f.m[last >> 6] |= 1ull << (last & 63);
            jpow -= j;
        }
        int sg = 1;
        res = majorana_mul(res, f, sg);
        sign *= sg;
    }
    phase = ipow(jpow) * (double)sign;
    return res;
}

inline PauliPolynomial jw_poly(const MajoranaPolynomial& a) {
    PauliPolynomial r;
    for (size_t i = 0; i < a.t.size(); ++i) {
        cplx ph;
        const PauliString ps = jw_string(a.t.keys[i], ph);
        r.t.add_compress(ps, a.t.vals[i] * ph);
    }
    return r;
}

inline MajoranaPolynomial majorana_poly(const PauliPolynomial& a) {
    MajoranaPolynomial r;
    for (size_t i = 0; i < a.t.size(); ++i) {
        cplx ph;
        const MajoranaString ms = majorana_of_pauli(a.t.keys[i], ph);
        r.t.add_compress(ms, a.t.vals[i] * ph);
    }
    return r;
}

// Image of a Majorana rotation exp(-i theta/2 i^r M): i^r M maps to s P with
// s = +-1 (a Hermitian generator maps to a Hermitian one), i.e. ROT(P, s theta).
inline RotGate jw_rotation(const MRotGate& g) {
    cplx ph;
    const PauliString ps = jw_string(g.ms, ph);
    const cplx h = (g.r ? cplx(0, 1) : cplx(1, 0)) * ph;
    if (h.imag() != 0.0 || std::abs(std::abs(h.real()) - 1.0) != 0.0) fail("internal error: non-Hermitian JW image");
    return RotGate{ps, h.real() > 0 ? g.theta : -g.theta};
}

}  // namespace ffh
