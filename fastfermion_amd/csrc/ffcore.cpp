// ffcore.cpp -- pybind11 module fastfermion_amd._ffcore: the drop-in for the
// reference's `ffcore.propagate` path.
//
// Mirrors the names, signatures, argument meaning and error behaviour of the
// reference bindings that the propagate path uses (src/python_bind.cpp:219-527,
// 579-916): PauliString / PauliPolynomial, the Clifford gates and ROT,
// MajoranaString / MajoranaPolynomial and MROT, QubitProductState / FockState,
// and the two `propagate` overloads (Pauli first, Majorana second).  The
// propagation itself runs on the GPU through the C ABI of libffamd.so
// (include/fastfermion_amd.h); there is no host fallback: without a usable HIP
// device `propagate` raises RuntimeError.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <mutex>
#include <optional>
#include <variant>

#include "fastfermion_amd.h"
#include "ffcore_types.h"

namespace py = pybind11;
using namespace ffh;

namespace {

// ---------------------------------------------------------------------------
// Pauli polynomial arithmetic (src/pauli/algebra.h:305-946)
// ---------------------------------------------------------------------------
PauliPolynomial pp_scalar(cplx v) {
    PauliPolynomial p;
    p.t.at(PauliString()) = v;
    return p;
}
PauliPolynomial pp_string(const PauliString& s, cplx v = 1.0) {
    PauliPolynomial p;
    p.t.at(s) = v;
    return p;
}
// PauliPolynomial(list of (index, op), coeff): product of the factors with the
// phase of repeated indices (src/pauli/algebra.h:313-327)
PauliPolynomial pp_from_indices(const std::vector<std::pair<int, char>>& s, cplx coeff) {
    PauliString u;
    int jpow = 0;
    for (const auto& f : s) {
        int j;
        u = pauli_mul(u, pauli_from_indices({f}), j);
        jpow += j;
    }
    PauliPolynomial p;
    p.t.at(u) = ipow(jpow) * coeff;
    return p;
}
void pp_iadd(PauliPolynomial& a, const PauliPolynomial& b, double sgn = 1.0) {
    for (size_t i = 0; i < b.t.size(); ++i) a.t.add_compress(b.t.keys[i], sgn * b.t.vals[i]);
}
void pp_iadd_scalar(PauliPolynomial& a, cplx v) {
    if (v != cplx(0, 0)) a.t.add_compress(PauliString(), v);
}
PauliPolynomial pp_mul(const PauliPolynomial& a, const PauliPolynomial& b) {
    PauliTerms nt;
    for (size_t i = 0; i < a.t.size(); ++i)
        for (size_t j = 0; j < b.t.size(); ++j) {
            int jp;
            const PauliString r = pauli_mul(a.t.keys[i], b.t.keys[j], jp);
            nt.at(r) += ipow(jp) * a.t.vals[i] * b.t.vals[j];
        }
    PauliPolynomial out;
    for (size_t i = 0; i < nt.size(); ++i)
        if (nt.vals[i] != cplx(0, 0)) out.t.at(nt.keys[i]) = nt.vals[i];
    return out;
}
PauliPolynomial pp_scale(const PauliPolynomial& a, cplx v) {
    PauliPolynomial c = a;
    for (auto& x : c.t.vals) x *= v;
    return c;
}
bool pp_eq(const PauliPolynomial& a, const PauliPolynomial& b) {
    PauliPolynomial c = a;
    pp_iadd(c, b, -1.0);
    for (auto& v : c.t.vals)
        if (v.real() != 0 || v.imag() != 0) return false;
    return true;
}
PauliPolynomial pp_commutator(const PauliPolynomial& a, const PauliPolynomial& b) {
    PauliPolynomial c;
    for (size_t i = 0; i < a.t.size(); ++i)
        for (size_t j = 0; j < b.t.size(); ++j)
            if (!a.t.keys[i].commutes(b.t.keys[j])) {
                int jp;
                const PauliString r = pauli_mul(a.t.keys[i], b.t.keys[j], jp);
                c.t.at(r) += 2.0 * ipow(jp) * a.t.vals[i] * b.t.vals[j];
            }
    return c;
}
bool pp_is_zero(const PauliPolynomial& p, double tol) {
    for (auto& v : p.t.vals)
        if (std::abs(v) > tol) return false;
    return true;
}

std::string fmt_float(double v) {
    std::ostringstream s;
    s << v;
    return s.str();
}
std::string fmt_complex(cplx z) {
    if (z.imag() == 0) return fmt_float(z.real());
    if (z.real() == 0) return fmt_float(z.imag()) + "j";
    if (z.imag() < 0) return "(" + fmt_float(z.real()) + "-" + fmt_float(-z.imag()) + "j)";
    return "(" + fmt_float(z.real()) + "+" + fmt_float(z.imag()) + "j)";
}
template <class T>
std::string fmt_poly(const T& t, int max_terms = 200) {
    if (t.size() == 0) return "0";
    std::string s;
    for (size_t i = 0; i < t.size(); ++i) {
        if ((int)i >= max_terms) {
            s += " + ... (" + std::to_string(t.size() - i) + " more terms)";
            break;
        }
        if (i) s += " + ";
        const bool ident = t.keys[i].is_identity();
        if (ident)
            s += fmt_complex(t.vals[i]);
        else if (t.vals[i] == cplx(1, 0))
            s += t.keys[i].compact();
        else
            s += fmt_complex(t.vals[i]) + " " + t.keys[i].compact();
    }
    return s;
}

// ---------------------------------------------------------------------------
// Majorana polynomial arithmetic (src/majorana_algebra.h:152-827)
// ---------------------------------------------------------------------------
MajoranaPolynomial mp_scalar(cplx v) {
    MajoranaPolynomial p;
    p.t.at(MajoranaString()) = v;
    return p;
}
MajoranaPolynomial mp_string(const MajoranaString& s, cplx v = 1.0) {
    MajoranaPolynomial p;
    p.t.at(s) = v;
    return p;
}
// MajoranaPolynomial(supp, coeff) with supp in any order (src/majorana_algebra.h:160-175)
MajoranaPolynomial mp_from_support(const std::vector<int>& supp, cplx coeff) {
    MajoranaString a;
    bool sgn = false;
    for (int i : supp) {
        if (i < 0 || i >= MAX_MAJORANAS)
            fail("Invalid MajoranaString (mode index " + std::to_string(i) + " invalid, should be between 0 and " +
                 std::to_string(MAX_MAJORANAS - 1) + ")");
        a.m[i >> 6] ^= 1ull << (i & 63);
        int above = 0;
        for (int q = i + 1; q < MAX_MAJORANAS; ++q) above += a.bit(q);
        sgn ^= (above & 1) != 0;
    }
    MajoranaPolynomial p;
    p.t.at(a) = (sgn ? -1.0 : 1.0) * coeff;
    return p;
}
void mp_iadd(MajoranaPolynomial& a, const MajoranaPolynomial& b, double sgn = 1.0) {
    for (size_t i = 0; i < b.t.size(); ++i) a.t.add_compress(b.t.keys[i], sgn * b.t.vals[i]);
}
void mp_iadd_scalar(MajoranaPolynomial& a, cplx v) {
    if (v != cplx(0, 0)) a.t.add_compress(MajoranaString(), v);
}
MajoranaPolynomial mp_mul(const MajoranaPolynomial& a, const MajoranaPolynomial& b) {
    MajoranaTerms nt;
    for (size_t i = 0; i < a.t.size(); ++i)
        for (size_t j = 0; j < b.t.size(); ++j) {
            int sg;
            const MajoranaString r = majorana_mul(a.t.keys[i], b.t.keys[j], sg);
            nt.at(r) += a.t.vals[i] * b.t.vals[j] * (double)sg;
        }
    MajoranaPolynomial out;
    for (size_t i = 0; i < nt.size(); ++i)
        if (nt.vals[i] != cplx(0, 0)) out.t.at(nt.keys[i]) = nt.vals[i];
    return out;
}
MajoranaPolynomial mp_scale(const MajoranaPolynomial& a, cplx v) {
    MajoranaPolynomial c = a;
    for (auto& x : c.t.vals) x *= v;
    return c;
}
bool mp_eq(const MajoranaPolynomial& a, const MajoranaPolynomial& b) {
    MajoranaPolynomial c = a;
    mp_iadd(c, b, -1.0);
    for (auto& v : c.t.vals)
        if (v.real() != 0 || v.imag() != 0) return false;
    return true;
}
MajoranaPolynomial mp_commutator(const MajoranaPolynomial& a, const MajoranaPolynomial& b) {
    MajoranaPolynomial c;
    for (size_t i = 0; i < a.t.size(); ++i)
        for (size_t j = 0; j < b.t.size(); ++j)
            if (!a.t.keys[i].commutes(b.t.keys[j])) {
                int sg;
                const MajoranaString r = majorana_mul(a.t.keys[i], b.t.keys[j], sg);
                c.t.at(r) += 2.0 * (double)sg * a.t.vals[i] * b.t.vals[j];
            }
    return c;
}
// m_S^dagger = (-1)^{d(d-1)/2} m_S (src/majorana_algebra.h:418-421)
MajoranaPolynomial mp_dagger(const MajoranaPolynomial& p) {
    MajoranaPolynomial b;
    for (size_t i = 0; i < p.t.size(); ++i) {
        const int d = p.t.keys[i].degree();
        const double s = ((d * (d - 1) / 2) % 2 == 0) ? 1.0 : -1.0;
        b.t.at(p.t.keys[i]) += s * std::conj(p.t.vals[i]);
    }
    return b;
}

// ---------------------------------------------------------------------------
// Gate conjugations on the host (gate.__call__): ROT::operator()
// (src/pauli/gates.h:234-252) and MROT::apply_inplace without filter
// (src/majorana_gates.h:67-103)
// ---------------------------------------------------------------------------
PauliPolynomial rot_apply(const RotGate& g, const PauliPolynomial& o) {
    PauliPolynomial obs = o;
    double c, s;
    trig(g.theta, c, s);
    const cplx isin(0, s);
    std::vector<std::pair<PauliString, cplx>> nt;
    for (size_t i = 0; i < obs.t.size(); ++i) {
        if (!obs.t.keys[i].commutes(g.ps)) {
            int jp;
            const PauliString r = pauli_mul(g.ps, obs.t.keys[i], jp);
            nt.emplace_back(r, obs.t.vals[i] * isin * ipow(jp));
            obs.t.vals[i] *= c;
        }
    }
    for (auto& kv : nt) obs.t.at(kv.first) += kv.second;
    return obs;
}

MajoranaPolynomial mrot_apply(const MRotGate& g, const MajoranaPolynomial& o) {
    MajoranaPolynomial p = o;
    double c, s;
    trig(g.theta, c, s);
    const cplx iirs = g.r ? cplx(-s, 0) : cplx(0, s);
    std::vector<std::pair<MajoranaString, cplx>> nt;
    for (size_t i = 0; i < p.t.size(); ++i) {
        if (!p.t.keys[i].commutes(g.ms)) {
            int sg;
            const MajoranaString r = majorana_mul(g.ms, p.t.keys[i], sg);
            nt.emplace_back(r, p.t.vals[i] * iirs * cplx((double)sg, 0));
            p.t.vals[i] *= c;
        }
    }
    for (auto& kv : nt) p.t.at(kv.first) += kv.second;
    return p;
}

// ---------------------------------------------------------------------------
// Device context (one per device, shared; calls are serialised)
// ---------------------------------------------------------------------------
std::mutex g_ctx_mu;
std::vector<ffx_ctx*> g_ctx;
ffx_profile g_last_profile{};
bool g_profile_on = false;

ffx_ctx* device_ctx(int device) {
    if (device < 0) throw py::value_error("device must be >= 0");
    if ((int)g_ctx.size() <= device) g_ctx.resize(device + 1, nullptr);
    if (!g_ctx[device]) {
        ffx_ctx* c = nullptr;
        const int rc = ffx_ctx_create(device, &c);
        if (rc != FFX_OK)
            throw std::runtime_error("fastfermion_amd: no usable HIP device " + std::to_string(device) +
                                     " (propagation runs only on the GPU; status " + std::to_string(rc) + ")");
        g_ctx[device] = c;
    }
    return g_ctx[device];
}

void check(ffx_ctx* c, int rc) {
    if (rc == FFX_OK) return;
    const std::string msg = ffx_last_error(c);
    if (rc == FFX_ENOMEM) throw std::bad_alloc();
    throw std::runtime_error("fastfermion_amd: " + msg);
}

int precision_code(const std::string& p) {
    if (p == "fp64") return FFX_FP64;
    if (p == "fp32") return FFX_FP32;
    if (p == "fp32c") return FFX_FP32C;
    throw py::value_error("precision must be 'fp64', 'fp32' or 'fp32c'");
}

using PauliGate = std::variant<GateH, GateS, GateCNOT, GateSWAP, GateCZ, RotGate>;
using PauliObservable = std::variant<PauliString, PauliPolynomial>;
using MajoranaObservable = std::variant<MajoranaString, MajoranaPolynomial>;

// single-gate Clifford conjugation of a string: (P', sign)
template <class G>
std::pair<PauliString, cplx> cliff_string(const G& g, int kind, int i, int j, const PauliString& a) {
    PauliString r = a;
    cplx c = 1.0;
    clifford_apply(kind, i, j, r, c);
    return {r, c};
}
PauliPolynomial cliff_poly(int kind, int i, int j, const PauliPolynomial& p) {
    PauliPolynomial res;
    for (size_t t = 0; t < p.t.size(); ++t) {
        PauliString r = p.t.keys[t];
        cplx c = 1.0;
        clifford_apply(kind, i, j, r, c);
        res.t.at(r) += c * p.t.vals[t];
    }
    return res;
}

void check_qubit(int q) {
    if (q < 0 || q >= MAX_QUBITS)
        fail("Invalid qubit index " + std::to_string(q) + " (should be between 0 and " + std::to_string(MAX_QUBITS - 1) + ")");
}

// ---------------------------------------------------------------------------
// propagate (Pauli): src/python_bind.cpp:392-425 -> src/pauli/propagate.h:297-493
// ---------------------------------------------------------------------------
struct PackedPauli {
    std::vector<ffx_pauli_gate> gates;
    std::vector<uint64_t> x, z;
    std::vector<double> re, im;
    int W = 1;
    bool complex = false;
};

PackedPauli pack_pauli(const std::vector<PauliGate>& circuit, const PauliPolynomial& obs) {
    PackedPauli pk;
    int bits = 1;
    for (const auto& g : circuit) {
        std::visit(
            [&](const auto& gg) {
                using T = std::decay_t<decltype(gg)>;
                if constexpr (std::is_same_v<T, RotGate>) {
                    bits = std::max(bits, gg.ps.bits_used());
                } else if constexpr (std::is_same_v<T, GateH> || std::is_same_v<T, GateS>) {
                    bits = std::max(bits, gg.i + 1);
                } else {
                    bits = std::max(bits, std::max(gg.i, gg.j) + 1);
                }
            },
            g);
    }
    for (const auto& k : obs.t.keys) bits = std::max(bits, k.bits_used());
    pk.W = (bits + 63) / 64;
    pk.gates.resize(circuit.size());
    for (size_t t = 0; t < circuit.size(); ++t) {
        ffx_pauli_gate& d = pk.gates[t];
        std::memset(&d, 0, sizeof d);
        std::visit(
            [&](const auto& gg) {
                using T = std::decay_t<decltype(gg)>;
                if constexpr (std::is_same_v<T, RotGate>) {
                    d.kind = FFX_ROT;
                    d.theta = gg.theta;
                    for (int w = 0; w < PW; ++w) {
                        d.x[w] = gg.ps.x[w];
                        d.z[w] = gg.ps.z[w];
                    }
                } else if constexpr (std::is_same_v<T, GateH>) {
                    d.kind = FFX_H;
                    d.q0 = gg.i;
                } else if constexpr (std::is_same_v<T, GateS>) {
                    d.kind = FFX_S;
                    d.q0 = gg.i;
                } else if constexpr (std::is_same_v<T, GateCNOT>) {
                    d.kind = FFX_CNOT;
                    d.q0 = gg.i;
                    d.q1 = gg.j;
                } else if constexpr (std::is_same_v<T, GateSWAP>) {
                    d.kind = FFX_SWAP;
                    d.q0 = gg.i;
                    d.q1 = gg.j;
                } else {
                    d.kind = FFX_CZ;
                    d.q0 = gg.i;
                    d.q1 = gg.j;
                }
            },
            circuit[t]);
    }
    const size_t n = obs.t.size();
    pk.x.resize(n * pk.W);
    pk.z.resize(n * pk.W);
    pk.re.resize(n);
    pk.im.resize(n);
    for (size_t i = 0; i < n; ++i) {
        for (int w = 0; w < pk.W; ++w) {
            pk.x[i * pk.W + w] = obs.t.keys[i].x[w];
            pk.z[i * pk.W + w] = obs.t.keys[i].z[w];
        }
        pk.re[i] = obs.t.vals[i].real();
        pk.im[i] = obs.t.vals[i].imag();
        if (pk.im[i] != 0.0) pk.complex = true;
    }
    return pk;
}

PauliPolynomial unpack_pauli(const ffx_terms& t) {
    PauliPolynomial p;
    const size_t n = (size_t)t.n_terms;
    const int W = t.n_words;
    p.t.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        PauliString k;
        for (int w = 0; w < W; ++w) {
            k.x[w] = t.x[i * W + w];
            k.z[w] = t.z[i * W + w];
        }
        p.t.at(k) = cplx(t.c_re[i], t.c_im[i]);
    }
    return p;
}

void warn_threads(int n_threads) {
    if (n_threads > 1)
        PyErr_WarnEx(PyExc_RuntimeWarning,
                     "n_threads > 1 ignored: fastfermion_amd propagates on the GPU (no OpenMP host path)", 1);
}

PauliPolynomial observable_pp(const PauliObservable& o) {
    return o.index() == 0 ? pp_string(std::get<0>(o)) : std::get<1>(o);
}

ffx_options make_opts(const std::optional<int>& maxdegree, const std::optional<double>& mincoeff, int topk,
                      int max_xweight, int xtrunc_period, bool batched, const std::string& precision) {
    ffx_options o{};
    o.maxdegree = maxdegree.has_value() ? maxdegree.value() : MAX_QUBITS;
    o.mincoeff = mincoeff.has_value() ? mincoeff.value() : 0.0;
    o.topk = topk;
    o.max_xweight = max_xweight;
    o.xtrunc_period = xtrunc_period;
    o.batched = batched ? 1 : 0;
    o.precision = precision_code(precision);
    return o;
}

PauliPolynomial propagate_pauli(const std::vector<PauliGate>& circuit, const PauliObservable& observable,
                                int n_threads, const std::optional<int>& maxdegree,
                                const std::optional<double>& mincoeff, int topk, int max_xweight, int xtrunc_period,
                                bool batched, const std::string& parallel, int device, const std::string& precision) {
    warn_threads(n_threads);
    (void)parallel;  // every strategy of the reference gives the same result; all run on the GPU
    const PauliPolynomial obs = observable_pp(observable);
    const ffx_options o = make_opts(maxdegree, mincoeff, topk, max_xweight, xtrunc_period, batched, precision);
    const PackedPauli pk = pack_pauli(circuit, obs);
    ffx_terms out{};
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    ffx_ctx* c = device_ctx(device);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = ffx_pauli_propagate(c, pk.gates.data(), (int64_t)pk.gates.size(), pk.W, pk.x.data(), pk.z.data(),
                                 pk.re.data(), pk.complex ? pk.im.data() : nullptr, (int64_t)pk.re.size(), &o, &out,
                                 nullptr, nullptr);
    }
    check(c, rc);
    if (g_profile_on) ffx_profile_get(c, &g_last_profile);
    PauliPolynomial res = unpack_pauli(out);
    ffx_terms_free(&out);
    return res;
}

// <b|O|b> on the device without downloading the terms (additive API).
cplx propagate_expectation_pauli(const std::vector<PauliGate>& circuit, const PauliObservable& observable,
                                 const std::vector<int>& up, const std::optional<int>& maxdegree,
                                 const std::optional<double>& mincoeff, int max_xweight, int xtrunc_period,
                                 bool batched, int device, const std::string& precision) {
    const PauliPolynomial obs = observable_pp(observable);
    const ffx_options o = make_opts(maxdegree, mincoeff, 0, max_xweight, xtrunc_period, batched, precision);
    PackedPauli pk = pack_pauli(circuit, obs);
    int bits = pk.W * 64;
    for (int q : up) {
        check_qubit(q);
        bits = std::max(bits, q + 1);
    }
    const int W = std::max(pk.W, (bits + 63) / 64);
    if (W != pk.W) {  // widen keys to the state's width
        std::vector<uint64_t> x(pk.re.size() * W, 0), z(pk.re.size() * W, 0);
        for (size_t i = 0; i < pk.re.size(); ++i)
            for (int w = 0; w < pk.W; ++w) {
                x[i * W + w] = pk.x[i * pk.W + w];
                z[i * W + w] = pk.z[i * pk.W + w];
            }
        pk.x.swap(x);
        pk.z.swap(z);
        pk.W = W;
    }
    uint64_t b[PW] = {0, 0, 0, 0};
    for (int q : up) b[q >> 6] |= 1ull << (q & 63);
    double ex[2] = {0, 0};
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    ffx_ctx* c = device_ctx(device);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = ffx_pauli_propagate(c, pk.gates.data(), (int64_t)pk.gates.size(), pk.W, pk.x.data(), pk.z.data(),
                                 pk.re.data(), pk.complex ? pk.im.data() : nullptr, (int64_t)pk.re.size(), &o, nullptr,
                                 b, ex);
    }
    check(c, rc);
    if (g_profile_on) ffx_profile_get(c, &g_last_profile);
    return cplx(ex[0], ex[1]);
}

// ---------------------------------------------------------------------------
// propagate (Majorana): src/python_bind.cpp:490-527 -> src/majorana_propagation.h:18-55
// ---------------------------------------------------------------------------
MajoranaPolynomial propagate_majorana(const std::vector<MRotGate>& circuit, const MajoranaObservable& observable,
                                      const std::optional<int>& maxdegree, const std::optional<double>& mincoeff,
                                      int device, const std::string& precision) {
    const MajoranaPolynomial obs = observable.index() == 0 ? mp_string(std::get<0>(observable)) : std::get<1>(observable);
    int bits = 1;
    for (const auto& g : circuit) bits = std::max(bits, g.ms.extent());
    for (const auto& k : obs.t.keys) bits = std::max(bits, k.extent());
    const int W = (bits + 63) / 64;
    std::vector<ffx_majorana_gate> gates(circuit.size());
    for (size_t t = 0; t < circuit.size(); ++t) {
        std::memset(&gates[t], 0, sizeof(ffx_majorana_gate));
        for (int w = 0; w < MW; ++w) gates[t].m[w] = circuit[t].ms.m[w];
        gates[t].theta = circuit[t].theta;
        gates[t].r = circuit[t].r ? 1 : 0;
    }
    const size_t n = obs.t.size();
    std::vector<uint64_t> m(n * W);
    std::vector<double> re(n), im(n);
    bool cx = false;
    for (size_t i = 0; i < n; ++i) {
        for (int w = 0; w < W; ++w) m[i * W + w] = obs.t.keys[i].m[w];
        re[i] = obs.t.vals[i].real();
        im[i] = obs.t.vals[i].imag();
        cx = cx || im[i] != 0.0;
    }
    const int md = maxdegree.has_value() ? maxdegree.value() : FFX_NO_DEGREE_LIMIT;
    const double mc = mincoeff.has_value() ? mincoeff.value() : 0.0;
    const int prec = precision_code(precision);
    ffx_terms out{};
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    ffx_ctx* c = device_ctx(device);
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = ffx_majorana_propagate(c, gates.data(), (int64_t)gates.size(), W, m.data(), re.data(),
                                    cx ? im.data() : nullptr, (int64_t)n, md, mc, prec, &out, nullptr, nullptr);
    }
    check(c, rc);
    if (g_profile_on) ffx_profile_get(c, &g_last_profile);
    MajoranaPolynomial res;
    res.t.reserve((size_t)out.n_terms);
    for (int64_t i = 0; i < out.n_terms; ++i) {
        MajoranaString k;
        for (int w = 0; w < out.n_words; ++w) k.m[w] = out.x[i * out.n_words + w];
        res.t.at(k) = cplx(out.c_re[i], out.c_im[i]);
    }
    ffx_terms_free(&out);
    return res;
}

// ---------------------------------------------------------------------------
// product-state readouts (src/qubitproductstate.h:73-88, src/fockstate.h:95-114)
// ---------------------------------------------------------------------------
struct QubitProductState {
    PauliString b;  // only x words used as the bit string
};
struct FockState {
    uint64_t b[PW] = {0, 0, 0, 0};  // mode occupations
};

cplx qps_string(const QubitProductState& s, const PauliString& a) {
    for (int w = 0; w < PW; ++w)
        if (a.x[w]) return 0.0;
    int par = 0;
    for (int w = 0; w < PW; ++w) par += popc(a.z[w] & s.b.x[w]);
    return (par & 1) ? -1.0 : 1.0;
}
cplx fock_string(const FockState& s, const MajoranaString& a) {
    int nj = 0, nb = 0;
    for (int w = 0; w < MW; ++w) {
        const uint64_t ev = a.m[w] & 0x5555555555555555ull;
        const uint64_t od = (a.m[w] >> 1) & 0x5555555555555555ull;
        if (ev != od) return 0.0;
        for (int bpos = 0; bpos < 64; bpos += 2)
            if ((ev >> bpos) & 1ull) {
                const int mode = (64 * w + bpos) / 2;
                ++nj;
                if ((s.b[mode >> 6] >> (mode & 63)) & 1ull) ++nb;
            }
    }
    return ipow(nj + 2 * nb);
}

std::vector<double> state_vec(const uint64_t* bits, int n, std::optional<int> nup) {
    if (n > 30) fail("Error: n is too large");
    int last = -1;
    int cnt = 0;
    for (int q = 0; q < PW * 64; ++q)
        if ((bits[q >> 6] >> (q & 63)) & 1ull) {
            last = q;
            ++cnt;
        }
    if (last >= n) fail("Error: n must be at least " + std::to_string(last + 1));
    uint64_t idx = 0;  // bit order reversed: qubit 0 is the most significant digit
    for (int q = 0; q < n; ++q)
        if ((bits[q >> 6] >> (q & 63)) & 1ull) idx |= 1ull << (n - 1 - q);
    if (!nup.has_value()) {
        std::vector<double> v(1ull << n, 0.0);
        v[idx] = 1.0;
        return v;
    }
    if (cnt != nup.value()) fail("Error: qubit state does not live in the supplied subspace");
    // rank of idx among n-bit strings of weight nup in increasing order
    std::vector<uint64_t> all;
    for (uint64_t v = 0; v < (1ull << n); ++v)
        if (popc(v) == nup.value()) all.push_back(v);
    std::vector<double> out(all.size(), 0.0);
    out[std::lower_bound(all.begin(), all.end(), idx) - all.begin()] = 1.0;
    return out;
}

}  // namespace

// ---------------------------------------------------------------------------
// Module (src/python_bind.cpp:579-916 for the types this path uses)
// ---------------------------------------------------------------------------
PYBIND11_MODULE(_ffcore, m, py::mod_gil_not_used()) {
    m.doc() = "MI355X propagation engine with the fastfermion ffcore interface for the propagate path";
    m.attr("MAX_QUBITS") = MAX_QUBITS;
    m.attr("FERMI_SYMBOL") = 'f';
    m.attr("MAJORANA_SYMBOL") = 'm';
    m.attr("DAGGER_SYMBOL") = '^';
    m.attr("__version__") = "0.1.0+mi355x";
    m.attr("has_openmp") = false;

    // ---------------- PauliString ----------------
    py::class_<PauliString> ps(m, "PauliString");
    ps.def(py::init<>())
        .def(py::init([](const std::string& s) { return pauli_from_string(s); }))
        .def(py::init([](const std::vector<std::pair<int, char>>& s) { return pauli_from_indices(s); }))
        .def("to_string", &PauliString::to_string, py::arg("n") = 0)
        .def("indices", &PauliString::indices)
        .def(
            "degree",
            [](const PauliString& a, const std::optional<char>& v) { return v.has_value() ? a.degree(v.value()) : a.degree(); },
            py::arg("v") = py::none())
        .def("extent", &PauliString::extent)
        .def("commutes", [](const PauliString& a, const PauliString& b) { return a.commutes(b); })
        .def("commutes", [](const PauliString& a, const PauliPolynomial& b) { return pp_is_zero(pp_commutator(pp_string(a), b), 0); })
        .def("commutator",
             [](const PauliString& a, const PauliString& b) { return pp_commutator(pp_string(a), pp_string(b)); })
        .def("commutator", [](const PauliString& a, const PauliPolynomial& b) { return pp_commutator(pp_string(a), b); })
        .def("__str__", &PauliString::compact)
        .def("__repr__", &PauliString::compact)
        .def("__hash__", [](const PauliString& a) { return a.hash(); })
        .def("__eq__", [](const PauliString& a, const PauliString& b) { return a == b; })
        .def("__eq__", [](const PauliString& a, cplx b) { return a.is_identity() && b == cplx(1, 0); })
        .def("__add__", [](const PauliString& a, cplx b) -> py::object {
            if (b == cplx(0, 0)) return py::cast(a);
            PauliPolynomial c = pp_string(a);
            pp_iadd_scalar(c, b);
            return py::cast(c);
        })
        .def("__add__", [](const PauliString& a, const PauliString& b) {
            PauliPolynomial c = pp_string(a);
            pp_iadd(c, pp_string(b));
            return c;
        })
        .def("__add__", [](const PauliString& a, const PauliPolynomial& b) {
            PauliPolynomial c = pp_string(a);
            pp_iadd(c, b);
            return c;
        })
        .def("__sub__", [](const PauliString& a, cplx b) -> py::object {
            if (b == cplx(0, 0)) return py::cast(a);
            PauliPolynomial c = pp_string(a);
            pp_iadd_scalar(c, -b);
            return py::cast(c);
        })
        .def("__sub__", [](const PauliString& a, const PauliString& b) {
            PauliPolynomial c = pp_string(a);
            pp_iadd(c, pp_string(b), -1.0);
            return c;
        })
        .def("__sub__", [](const PauliString& a, const PauliPolynomial& b) {
            PauliPolynomial c = pp_string(a);
            pp_iadd(c, b, -1.0);
            return c;
        })
        .def("__mul__", [](const PauliString& a, cplx b) { return pp_string(a, b); })
        .def("__mul__", [](const PauliString& a, const PauliString& b) { return pp_mul(pp_string(a), pp_string(b)); })
        .def("__mul__", [](const PauliString& a, const PauliPolynomial& b) { return pp_mul(pp_string(a), b); })
        .def("__rmul__", [](const PauliString& a, cplx b) { return pp_string(a, b); })
        .def("__radd__", [](const PauliString& a, cplx b) {
            PauliPolynomial c = pp_scalar(b);
            if (b == cplx(0, 0)) c = PauliPolynomial();
            pp_iadd(c, pp_string(a));
            return c;
        })
        .def("__rsub__", [](const PauliString& a, cplx b) {
            PauliPolynomial c = b == cplx(0, 0) ? PauliPolynomial() : pp_scalar(b);
            pp_iadd(c, pp_string(a), -1.0);
            return c;
        })
        .def("__neg__", [](const PauliString& a) { return pp_string(a, -1.0); });

    // ---------------- PauliPolynomial ----------------
    py::class_<PauliPolynomial> pp(m, "PauliPolynomial");
    pp.def(py::init<>())
        .def(py::init([](cplx v) { return pp_scalar(v); }))
        .def(py::init([](const PauliString& s) { return pp_string(s); }))
        .def(py::init([](const std::vector<std::pair<int, char>>& s) { return pp_from_indices(s, 1.0); }))
        .def(py::init([](const std::vector<std::pair<int, char>>& s, cplx c) { return pp_from_indices(s, c); }))
        .def(py::init([](const PauliPolynomial& p) { return p; }))
        .def(
            "degree",
            [](const PauliPolynomial& a, const std::optional<char>& v) {
                int d = 0;
                for (size_t i = 0; i < a.t.size(); ++i)
                    if (a.t.vals[i] != cplx(0, 0))
                        d = std::max(d, v.has_value() ? a.t.keys[i].degree(v.value()) : a.t.keys[i].degree());
                return d;
            },
            py::arg("v") = py::none())
        .def("dagger", [](const PauliPolynomial& a) {
            PauliPolynomial b;
            for (size_t i = 0; i < a.t.size(); ++i) b.t.at(a.t.keys[i]) = std::conj(a.t.vals[i]);
            return b;
        })
        .def("coefficient", [](const PauliPolynomial& a, const PauliString& s) {
            const cplx* v = a.t.find(s);
            return v ? *v : cplx(0, 0);
        })
        .def("coefficient", [](const PauliPolynomial& a, const std::vector<std::pair<int, char>>& s) {
            const cplx* v = a.t.find(pauli_from_indices(s));
            return v ? *v : cplx(0, 0);
        })
        .def("coefficient", [](const PauliPolynomial& a, const std::string& s) {
            const cplx* v = a.t.find(pauli_from_string(s));
            return v ? *v : cplx(0, 0);
        })
        .def("support", [](const PauliPolynomial& a) {
            PauliString u;
            for (size_t i = 0; i < a.t.size(); ++i)
                if (a.t.vals[i] != cplx(0, 0))
                    for (int w = 0; w < PW; ++w) u.x[w] |= a.t.keys[i].x[w] | a.t.keys[i].z[w];
            return u.support();
        })
        // sum of the coefficients of I/Z-only strings (src/pauli/algebra.h:549-557)
        .def("overlapwithzero", [](const PauliPolynomial& a) {
            cplx r = 0;
            for (size_t i = 0; i < a.t.size(); ++i) {
                const PauliString& k = a.t.keys[i];
                if (!(k.x[0] | k.x[1] | k.x[2] | k.x[3])) r += a.t.vals[i];
            }
            return r;
        })
        .def_property_readonly("terms", [](const PauliPolynomial& a) {
            py::dict d;
            for (size_t i = 0; i < a.t.size(); ++i) d[py::cast(a.t.keys[i])] = py::cast(a.t.vals[i]);
            return d;
        })
        .def("__len__", [](const PauliPolynomial& a) { return a.t.size(); })
        .def("__str__", [](const PauliPolynomial& a) { return fmt_poly(a.t); })
        .def("__repr__", [](const PauliPolynomial& a) { return fmt_poly(a.t); })
        .def("extent", [](const PauliPolynomial& a) {
            int l = 0;
            for (size_t i = 0; i < a.t.size(); ++i)
                if (a.t.vals[i] != cplx(0, 0)) l = std::max(l, a.t.keys[i].extent());
            return l;
        })
        .def("compress", [](const PauliPolynomial& a, double thr) {
            PauliPolynomial b;
            for (size_t i = 0; i < a.t.size(); ++i)
                if (std::abs(a.t.vals[i]) > thr) b.t.at(a.t.keys[i]) = a.t.vals[i];
            return b;
        }, py::arg("threshold") = 1e-8)
        .def("truncate", [](const PauliPolynomial& a, int k) {
            PauliPolynomial b;
            for (size_t i = 0; i < a.t.size(); ++i)
                if (a.t.keys[i].degree() <= k) b.t.at(a.t.keys[i]) = a.t.vals[i];
            return b;
        })
        .def("norm", [](const PauliPolynomial& a, const std::variant<int, std::string>& v) {
            if (v.index() == 1) {
                if (std::get<1>(v) != "inf") fail("Invalid argument");
                double r = 0;
                for (auto& x : a.t.vals) r = std::max(r, std::abs(x));
                return r;
            }
            const int p = std::get<0>(v);
            double r = 0;
            for (auto& x : a.t.vals) {
                if (p == 0) r += 1;
                else if (p == 1) r += std::abs(x);
                else if (p == 2) r += std::abs(x) * std::abs(x);
                else fail("Invalid value of p (allowed values are p=0,p=1,p=2)");
            }
            return p == 2 ? std::sqrt(r) : r;
        }, py::arg("v") = 1)
        .def("commutes", [](const PauliPolynomial& a, const PauliString& b) { return pp_is_zero(pp_commutator(a, pp_string(b)), 0); })
        .def("commutes", [](const PauliPolynomial& a, const PauliPolynomial& b) { return pp_is_zero(pp_commutator(a, b), 0); })
        .def("commutator", [](const PauliPolynomial& a, const PauliString& b) { return pp_commutator(a, pp_string(b)); })
        .def("commutator", [](const PauliPolynomial& a, const PauliPolynomial& b) { return pp_commutator(a, b); })
        // device-friendly export / import of large polynomials (no dict)
        .def("to_arrays", [](const PauliPolynomial& a) {
            const size_t n = a.t.size();
            py::array_t<uint64_t> x({n, (size_t)PW}), z({n, (size_t)PW});
            py::array_t<cplx> c(n);
            auto xm = x.mutable_unchecked<2>();
            auto zm = z.mutable_unchecked<2>();
            auto cm = c.mutable_unchecked<1>();
            for (size_t i = 0; i < n; ++i) {
                for (int w = 0; w < PW; ++w) {
                    xm(i, w) = a.t.keys[i].x[w];
                    zm(i, w) = a.t.keys[i].z[w];
                }
                cm(i) = a.t.vals[i];
            }
            return py::make_tuple(x, z, c);
        })
        .def_static("from_arrays", [](py::array_t<uint64_t, py::array::c_style | py::array::forcecast> x,
                                      py::array_t<uint64_t, py::array::c_style | py::array::forcecast> z,
                                      py::array_t<cplx, py::array::c_style | py::array::forcecast> c) {
            if (x.ndim() != 2 || z.ndim() != 2 || c.ndim() != 1 || x.shape(0) != c.shape(0) || z.shape(0) != c.shape(0) ||
                x.shape(1) != z.shape(1) || x.shape(1) > PW)
                throw py::value_error("expected x, z of shape (n, W<=4) and c of shape (n,)");
            const size_t n = (size_t)c.shape(0), W = (size_t)x.shape(1);
            PauliPolynomial p;
            p.t.reserve(n);
            auto xr = x.unchecked<2>();
            auto zr = z.unchecked<2>();
            auto cr = c.unchecked<1>();
            for (size_t i = 0; i < n; ++i) {
                PauliString k;
                for (size_t w = 0; w < W; ++w) {
                    k.x[w] = xr(i, w);
                    k.z[w] = zr(i, w);
                }
                p.t.at(k) += cr(i);
            }
            return p;
        });
    // arithmetic (src/python_bind.cpp:86-124)
    pp.def("__iadd__", [](PauliPolynomial& a, cplx b) -> PauliPolynomial& { pp_iadd_scalar(a, b); return a; })
        .def("__iadd__", [](PauliPolynomial& a, const PauliString& b) -> PauliPolynomial& { pp_iadd(a, pp_string(b)); return a; })
        .def("__iadd__", [](PauliPolynomial& a, const PauliPolynomial& b) -> PauliPolynomial& { pp_iadd(a, b); return a; })
        .def("__isub__", [](PauliPolynomial& a, cplx b) -> PauliPolynomial& { pp_iadd_scalar(a, -b); return a; })
        .def("__isub__", [](PauliPolynomial& a, const PauliString& b) -> PauliPolynomial& { pp_iadd(a, pp_string(b), -1.0); return a; })
        .def("__isub__", [](PauliPolynomial& a, const PauliPolynomial& b) -> PauliPolynomial& { pp_iadd(a, b, -1.0); return a; })
        .def("__imul__", [](PauliPolynomial& a, cplx b) -> PauliPolynomial& { a = pp_scale(a, b); return a; })
        .def("__imul__", [](PauliPolynomial& a, const PauliString& b) -> PauliPolynomial& { a = pp_mul(a, pp_string(b)); return a; })
        .def("__imul__", [](PauliPolynomial& a, const PauliPolynomial& b) -> PauliPolynomial& { a = pp_mul(a, b); return a; })
        .def("__itruediv__", [](PauliPolynomial& a, cplx b) -> PauliPolynomial& { for (auto& v : a.t.vals) v /= b; return a; })
        .def("__neg__", [](const PauliPolynomial& a) { return pp_scale(a, -1.0); })
        .def("__add__", [](const PauliPolynomial& a, cplx b) { PauliPolynomial c = a; pp_iadd_scalar(c, b); return c; })
        .def("__add__", [](const PauliPolynomial& a, const PauliString& b) { PauliPolynomial c = a; pp_iadd(c, pp_string(b)); return c; })
        .def("__add__", [](const PauliPolynomial& a, const PauliPolynomial& b) { PauliPolynomial c = a; pp_iadd(c, b); return c; })
        .def("__sub__", [](const PauliPolynomial& a, cplx b) { PauliPolynomial c = a; pp_iadd_scalar(c, -b); return c; })
        .def("__sub__", [](const PauliPolynomial& a, const PauliString& b) { PauliPolynomial c = a; pp_iadd(c, pp_string(b), -1.0); return c; })
        .def("__sub__", [](const PauliPolynomial& a, const PauliPolynomial& b) { PauliPolynomial c = a; pp_iadd(c, b, -1.0); return c; })
        .def("__mul__", [](const PauliPolynomial& a, cplx b) { return pp_scale(a, b); })
        .def("__mul__", [](const PauliPolynomial& a, const PauliString& b) { return pp_mul(a, pp_string(b)); })
        .def("__mul__", [](const PauliPolynomial& a, const PauliPolynomial& b) { return pp_mul(a, b); })
        .def("__truediv__", [](const PauliPolynomial& a, cplx b) { PauliPolynomial c = a; for (auto& v : c.t.vals) v /= b; return c; })
        .def("__radd__", [](const PauliPolynomial& a, cplx b) { PauliPolynomial c = pp_scalar(b); pp_iadd(c, a); return c; })
        .def("__rsub__", [](const PauliPolynomial& a, cplx b) { PauliPolynomial c = pp_scalar(b); pp_iadd(c, a, -1.0); return c; })
        .def("__rmul__", [](const PauliPolynomial& a, cplx b) { return pp_scale(a, b); })
        .def("__eq__", [](const PauliPolynomial& a, const PauliPolynomial& b) { return pp_eq(a, b); })
        .def("__eq__", [](const PauliPolynomial& a, const PauliString& b) { return pp_eq(a, pp_string(b)); })
        .def("__eq__", [](const PauliPolynomial& a, cplx b) { return pp_eq(a, pp_scalar(b)); })
        .def("__ne__", [](const PauliPolynomial& a, const PauliPolynomial& b) { return !pp_eq(a, b); });

    // ---------------- MajoranaString / MajoranaPolynomial ----------------
    py::class_<MajoranaString> ms(m, "MajoranaString");
    py::class_<MajoranaPolynomial> mpc(m, "MajoranaPolynomial");
    ms.def(py::init<>())
        .def(py::init([](const std::vector<int>& s) { return majorana_from_support(s); }))
        .def("extent", &MajoranaString::extent)
        .def("degree", &MajoranaString::degree)
        .def("indices", &MajoranaString::support)
        .def("is_hermitian", &MajoranaString::is_hermitian)
        .def("commutes", [](const MajoranaString& a, const MajoranaString& b) { return a.commutes(b); })
        .def("commutes", [](const MajoranaString& a, const MajoranaPolynomial& b) {
            for (auto& v : mp_commutator(mp_string(a), b).t.vals)
                if (v != cplx(0, 0)) return false;
            return true;
        })
        .def("commutator", [](const MajoranaString& a, const MajoranaString& b) { return mp_commutator(mp_string(a), mp_string(b)); })
        .def("commutator", [](const MajoranaString& a, const MajoranaPolynomial& b) { return mp_commutator(mp_string(a), b); })
        .def("dagger", [](const MajoranaString& a) { return mp_dagger(mp_string(a)); })
        .def("__hash__", [](const MajoranaString& a) { return a.hash(); })
        .def("__eq__", [](const MajoranaString& a, const MajoranaString& b) { return a == b; })
        .def("__eq__", [](const MajoranaString& a, cplx b) { return a.is_identity() && b == cplx(1, 0); })
        .def("__neg__", [](const MajoranaString& a) { return mp_string(a, -1.0); })
        .def("__add__", [](const MajoranaString& a, cplx b) { MajoranaPolynomial c = mp_string(a); mp_iadd_scalar(c, b); return c; })
        .def("__add__", [](const MajoranaString& a, const MajoranaString& b) { MajoranaPolynomial c = mp_string(a); mp_iadd(c, mp_string(b)); return c; })
        .def("__add__", [](const MajoranaString& a, const MajoranaPolynomial& b) { MajoranaPolynomial c = mp_string(a); mp_iadd(c, b); return c; })
        .def("__sub__", [](const MajoranaString& a, cplx b) { MajoranaPolynomial c = mp_string(a); mp_iadd_scalar(c, -b); return c; })
        .def("__sub__", [](const MajoranaString& a, const MajoranaString& b) { MajoranaPolynomial c = mp_string(a); mp_iadd(c, mp_string(b), -1.0); return c; })
        .def("__sub__", [](const MajoranaString& a, const MajoranaPolynomial& b) { MajoranaPolynomial c = mp_string(a); mp_iadd(c, b, -1.0); return c; })
        .def("__mul__", [](const MajoranaString& a, cplx b) { return mp_string(a, b); })
        .def("__mul__", [](const MajoranaString& a, const MajoranaString& b) { return mp_mul(mp_string(a), mp_string(b)); })
        .def("__mul__", [](const MajoranaString& a, const MajoranaPolynomial& b) { return mp_mul(mp_string(a), b); })
        .def("__rmul__", [](const MajoranaString& a, cplx b) { return mp_string(a, b); })
        .def("__radd__", [](const MajoranaString& a, cplx b) { MajoranaPolynomial c = mp_scalar(b); mp_iadd(c, mp_string(a)); return c; })
        .def("__rsub__", [](const MajoranaString& a, cplx b) { MajoranaPolynomial c = mp_scalar(b); mp_iadd(c, mp_string(a), -1.0); return c; })
        .def("__str__", &MajoranaString::compact)
        .def("__repr__", &MajoranaString::compact);

    mpc.def(py::init<>())
        .def(py::init([](cplx v) { return mp_scalar(v); }))
        .def(py::init([](const MajoranaString& s) { return mp_string(s); }))
        .def(py::init([](const MajoranaPolynomial& p) { return p; }))
        .def(py::init([](const std::vector<int>& s) { return mp_from_support(s, 1.0); }))
        .def(py::init([](const std::vector<int>& s, cplx c) { return mp_from_support(s, c); }))
        .def("degree", [](const MajoranaPolynomial& a) {
            int d = 0;
            for (size_t i = 0; i < a.t.size(); ++i)
                if (a.t.vals[i] != cplx(0, 0)) d = std::max(d, a.t.keys[i].degree());
            return d;
        })
        .def("dagger", &mp_dagger)
        .def("coefficient", [](const MajoranaPolynomial& a, const MajoranaString& s) {
            const cplx* v = a.t.find(s);
            return v ? *v : cplx(0, 0);
        })
        .def("coefficient", [](const MajoranaPolynomial& a, const std::vector<int>& s) {
            const cplx* v = a.t.find(majorana_from_support(s));
            return v ? *v : cplx(0, 0);
        })
        .def("support", [](const MajoranaPolynomial& a) {
            MajoranaString u;
            for (size_t i = 0; i < a.t.size(); ++i)
                if (a.t.vals[i] != cplx(0, 0))
                    for (int w = 0; w < MW; ++w) u.m[w] |= a.t.keys[i].m[w];
            return u.support();
        })
        // vacuum expectation (src/majorana_algebra.h:386-402)
        .def("overlapwithvacuum", [](const MajoranaPolynomial& a) {
            cplx r = 0;
            FockState vac;
            for (size_t i = 0; i < a.t.size(); ++i) r += a.t.vals[i] * fock_string(vac, a.t.keys[i]);
            return r;
        })
        .def_property_readonly("terms", [](const MajoranaPolynomial& a) {
            py::dict d;
            for (size_t i = 0; i < a.t.size(); ++i) d[py::cast(a.t.keys[i])] = py::cast(a.t.vals[i]);
            return d;
        })
        .def("__len__", [](const MajoranaPolynomial& a) { return a.t.size(); })
        .def("__str__", [](const MajoranaPolynomial& a) { return fmt_poly(a.t); })
        .def("__repr__", [](const MajoranaPolynomial& a) { return fmt_poly(a.t); })
        .def("extent", [](const MajoranaPolynomial& a) {
            int l = 0;
            for (size_t i = 0; i < a.t.size(); ++i)
                if (a.t.vals[i] != cplx(0, 0)) l = std::max(l, a.t.keys[i].extent());
            return l;
        })
        .def("compress", [](const MajoranaPolynomial& a, double thr) {
            MajoranaPolynomial b;
            for (size_t i = 0; i < a.t.size(); ++i)
                if (std::abs(a.t.vals[i]) > thr) b.t.at(a.t.keys[i]) = a.t.vals[i];
            return b;
        }, py::arg("threshold") = 1e-8)
        .def("truncate", [](const MajoranaPolynomial& a, int k) {
            MajoranaPolynomial b;
            for (size_t i = 0; i < a.t.size(); ++i)
                if (a.t.keys[i].degree() <= k) b.t.at(a.t.keys[i]) = a.t.vals[i];
            return b;
        })
        .def("norm", [](const MajoranaPolynomial& a, const std::variant<int, std::string>& v) {
            if (v.index() == 1) {
                if (std::get<1>(v) != "inf") fail("Invalid argument");
                double r = 0;
                for (auto& x : a.t.vals) r = std::max(r, std::abs(x));
                return r;
            }
            const int p = std::get<0>(v);
            double r = 0;
            for (auto& x : a.t.vals) {
                if (p == 0) r += 1;
                else if (p == 1) r += std::abs(x);
                else if (p == 2) r += std::abs(x) * std::abs(x);
                else fail("Invalid value of p (allowed values are p=0,p=1,p=2)");
            }
            return p == 2 ? std::sqrt(r) : r;
        }, py::arg("v") = 1)
        .def("commutator", [](const MajoranaPolynomial& a, const MajoranaPolynomial& b) { return mp_commutator(a, b); })
        .def("__iadd__", [](MajoranaPolynomial& a, cplx b) -> MajoranaPolynomial& { mp_iadd_scalar(a, b); return a; })
        .def("__iadd__", [](MajoranaPolynomial& a, const MajoranaString& b) -> MajoranaPolynomial& { mp_iadd(a, mp_string(b)); return a; })
        .def("__iadd__", [](MajoranaPolynomial& a, const MajoranaPolynomial& b) -> MajoranaPolynomial& { mp_iadd(a, b); return a; })
        .def("__isub__", [](MajoranaPolynomial& a, cplx b) -> MajoranaPolynomial& { mp_iadd_scalar(a, -b); return a; })
        .def("__isub__", [](MajoranaPolynomial& a, const MajoranaPolynomial& b) -> MajoranaPolynomial& { mp_iadd(a, b, -1.0); return a; })
        .def("__imul__", [](MajoranaPolynomial& a, cplx b) -> MajoranaPolynomial& { a = mp_scale(a, b); return a; })
        .def("__imul__", [](MajoranaPolynomial& a, const MajoranaPolynomial& b) -> MajoranaPolynomial& { a = mp_mul(a, b); return a; })
        .def("__neg__", [](const MajoranaPolynomial& a) { return mp_scale(a, -1.0); })
        .def("__add__", [](const MajoranaPolynomial& a, cplx b) { MajoranaPolynomial c = a; mp_iadd_scalar(c, b); return c; })
        .def("__add__", [](const MajoranaPolynomial& a, const MajoranaString& b) { MajoranaPolynomial c = a; mp_iadd(c, mp_string(b)); return c; })
        .def("__add__", [](const MajoranaPolynomial& a, const MajoranaPolynomial& b) { MajoranaPolynomial c = a; mp_iadd(c, b); return c; })
        .def("__sub__", [](const MajoranaPolynomial& a, cplx b) { MajoranaPolynomial c = a; mp_iadd_scalar(c, -b); return c; })
        .def("__sub__", [](const MajoranaPolynomial& a, const MajoranaString& b) { MajoranaPolynomial c = a; mp_iadd(c, mp_string(b), -1.0); return c; })
        .def("__sub__", [](const MajoranaPolynomial& a, const MajoranaPolynomial& b) { MajoranaPolynomial c = a; mp_iadd(c, b, -1.0); return c; })
        .def("__mul__", [](const MajoranaPolynomial& a, cplx b) { return mp_scale(a, b); })
        .def("__mul__", [](const MajoranaPolynomial& a, const MajoranaString& b) { return mp_mul(a, mp_string(b)); })
        .def("__mul__", [](const MajoranaPolynomial& a, const MajoranaPolynomial& b) { return mp_mul(a, b); })
        .def("__truediv__", [](const MajoranaPolynomial& a, cplx b) { MajoranaPolynomial c = a; for (auto& v : c.t.vals) v /= b; return c; })
        .def("__radd__", [](const MajoranaPolynomial& a, cplx b) { MajoranaPolynomial c = mp_scalar(b); mp_iadd(c, a); return c; })
        .def("__rsub__", [](const MajoranaPolynomial& a, cplx b) { MajoranaPolynomial c = mp_scalar(b); mp_iadd(c, a, -1.0); return c; })
        .def("__rmul__", [](const MajoranaPolynomial& a, cplx b) { return mp_scale(a, b); })
        .def("__eq__", [](const MajoranaPolynomial& a, const MajoranaPolynomial& b) { return mp_eq(a, b); })
        .def("__ne__", [](const MajoranaPolynomial& a, const MajoranaPolynomial& b) { return !mp_eq(a, b); });

    // ---------------- states ----------------
    py::class_<QubitProductState>(m, "QubitProductState")
        .def(py::init<>())
        .def(py::init([](uint64_t b) { QubitProductState s; s.b.x[0] = b; return s; }))
        .def(py::init([](const std::vector<int>& up) {
            QubitProductState s;
            for (int q : up) { check_qubit(q); s.b.x[q >> 6] |= 1ull << (q & 63); }
            return s;
        }), py::arg("up") = std::vector<int>{})
        .def("__call__", [](const QubitProductState& s, const PauliString& a) { return qps_string(s, a); })
        .def("__call__", [](const QubitProductState& s, const PauliPolynomial& p) {
            cplx r = 0;
            for (size_t i = 0; i < p.t.size(); ++i) r += p.t.vals[i] * qps_string(s, p.t.keys[i]);
            return r;
        })
        .def("up", [](const QubitProductState& s) {
            std::vector<int> r;
            for (int q = 0; q < MAX_QUBITS; ++q) if (s.b.xbit(q)) r.push_back(q);
            return r;
        })
        .def("vec", [](const QubitProductState& s, int n, std::optional<int> nup) { return state_vec(s.b.x, n, nup); },
             py::arg("n"), py::arg("nocc") = py::none());

    py::class_<FockState>(m, "FockState")
        .def(py::init<>())
        .def(py::init([](uint64_t b) { FockState s; s.b[0] = b; return s; }))
        .def(py::init([](const std::vector<int>& occ) {
            FockState s;
            for (int q : occ) { check_qubit(q); s.b[q >> 6] |= 1ull << (q & 63); }
            return s;
        }), py::arg("occ") = std::vector<int>{})
        .def("__call__", [](const FockState& s, const MajoranaString& a) { return fock_string(s, a); })
        .def("__call__", [](const FockState& s, const MajoranaPolynomial& p) {
            cplx r = 0;
            for (size_t i = 0; i < p.t.size(); ++i) r += p.t.vals[i] * fock_string(s, p.t.keys[i]);
            return r;
        })
        .def("occ", [](const FockState& s) {
            std::vector<int> r;
            for (int q = 0; q < MAX_QUBITS; ++q) if ((s.b[q >> 6] >> (q & 63)) & 1ull) r.push_back(q);
            return r;
        })
        .def("vec", [](const FockState& s, int n, std::optional<int> nocc) { return state_vec(s.b, n, nocc); },
             py::arg("n"), py::arg("nocc") = py::none());

    // ---------------- gates (src/python_bind.cpp:264-390) ----------------
    py::class_<GateH>(m, "H")
        .def(py::init([](int i) { check_qubit(i); return GateH{i}; }))
        .def("__call__", [](const GateH& g, const PauliString& a) { return cliff_string(g, kH, g.i, 0, a); })
        .def("__call__", [](const GateH& g, const PauliPolynomial& p) { return cliff_poly(kH, g.i, 0, p); })
        .def_property_readonly("qubits", [](const GateH& g) { return std::vector<int>{g.i}; })
        .def("aspoly", [](const GateH& g) {
            PauliPolynomial p;
            PauliString x, z;
            x.set(g.i, 'X');
            z.set(g.i, 'Z');
            p.t.at(x) = 1 / std::sqrt(2.0);
            p.t.at(z) = 1 / std::sqrt(2.0);
            return p;
        })
        .def("__repr__", [](const GateH& g) { return "H(" + std::to_string(g.i) + ")"; })
        .def("__str__", [](const GateH& g) { return "H(" + std::to_string(g.i) + ")"; });
    py::class_<GateS>(m, "S")
        .def(py::init([](int i) { check_qubit(i); return GateS{i}; }))
        .def("__call__", [](const GateS& g, const PauliString& a) { return cliff_string(g, kS, g.i, 0, a); })
        .def("__call__", [](const GateS& g, const PauliPolynomial& p) { return cliff_poly(kS, g.i, 0, p); })
        .def_property_readonly("qubits", [](const GateS& g) { return std::vector<int>{g.i}; })
        .def("aspoly", [](const GateS& g) {
            PauliPolynomial p;
            PauliString z;
            z.set(g.i, 'Z');
            p.t.at(PauliString()) = cplx(.5, .5);
            p.t.at(z) = cplx(.5, -.5);
            return p;
        })
        .def("__repr__", [](const GateS& g) { return "S(" + std::to_string(g.i) + ")"; })
        .def("__str__", [](const GateS& g) { return "S(" + std::to_string(g.i) + ")"; });
    auto two = [](int i, int j) {
        check_qubit(i);
        check_qubit(j);
        if (i == j) fail("two-qubit gate needs two distinct qubits");
    };
    py::class_<GateCNOT>(m, "CNOT")
        .def(py::init([two](int i, int j) { two(i, j); return GateCNOT{i, j}; }))
        .def("__call__", [](const GateCNOT& g, const PauliString& a) { return cliff_string(g, kCNOT, g.i, g.j, a); })
        .def("__call__", [](const GateCNOT& g, const PauliPolynomial& p) { return cliff_poly(kCNOT, g.i, g.j, p); })
        .def_property_readonly("qubits", [](const GateCNOT& g) { return std::vector<int>{g.i, g.j}; })
        .def("aspoly", [](const GateCNOT& g) {
            PauliPolynomial p;
            PauliString zi, xj, zx;
            zi.set(g.i, 'Z');
            xj.set(g.j, 'X');
            zx.set(g.i, 'Z');
            zx.set(g.j, 'X');
            p.t.at(PauliString()) = 0.5;
            p.t.at(zi) = 0.5;
            p.t.at(xj) = 0.5;
            p.t.at(zx) = -0.5;
            return p;
        })
        .def("__repr__", [](const GateCNOT& g) { return "CNOT(" + std::to_string(g.i) + "," + std::to_string(g.j) + ")"; })
        .def("__str__", [](const GateCNOT& g) { return "CNOT(" + std::to_string(g.i) + "," + std::to_string(g.j) + ")"; });
    py::class_<GateSWAP>(m, "SWAP")
        .def(py::init([two](int i, int j) { two(i, j); return GateSWAP{i, j}; }))
        .def("__call__", [](const GateSWAP& g, const PauliString& a) { return cliff_string(g, kSWAP, g.i, g.j, a); })
        .def("__call__", [](const GateSWAP& g, const PauliPolynomial& p) { return cliff_poly(kSWAP, g.i, g.j, p); })
        .def_property_readonly("qubits", [](const GateSWAP& g) { return std::vector<int>{g.i, g.j}; })
        .def("aspoly", [](const GateSWAP& g) {
            PauliPolynomial p;
            PauliString xx, yy, zz;
            xx.set(g.i, 'X'); xx.set(g.j, 'X');
            yy.set(g.i, 'Y'); yy.set(g.j, 'Y');
            zz.set(g.i, 'Z'); zz.set(g.j, 'Z');
            p.t.at(PauliString()) = 0.5;
            p.t.at(xx) = 0.5;
            p.t.at(yy) = 0.5;
            p.t.at(zz) = 0.5;
            return p;
        })
        .def("__repr__", [](const GateSWAP& g) { return "SWAP(" + std::to_string(g.i) + "," + std::to_string(g.j) + ")"; })
        .def("__str__", [](const GateSWAP& g) { return "SWAP(" + std::to_string(g.i) + "," + std::to_string(g.j) + ")"; });
    py::class_<GateCZ>(m, "CZ")
        .def(py::init([two](int i, int j) { two(i, j); return GateCZ{i, j}; }))
        .def("__call__", [](const GateCZ& g, const PauliString& a) { return cliff_string(g, kCZ, g.i, g.j, a); })
        .def("__call__", [](const GateCZ& g, const PauliPolynomial& p) { return cliff_poly(kCZ, g.i, g.j, p); })
        .def_property_readonly("qubits", [](const GateCZ& g) { return std::vector<int>{g.i, g.j}; })
        .def("aspoly", [](const GateCZ& g) {
            PauliPolynomial p;
            PauliString zi, zj, zz;
            zi.set(g.i, 'Z');
            zj.set(g.j, 'Z');
            zz.set(g.i, 'Z');
            zz.set(g.j, 'Z');
            p.t.at(PauliString()) = 0.5;
            p.t.at(zi) = 0.5;
            p.t.at(zj) = 0.5;
            p.t.at(zz) = -0.5;
            return p;
        })
        .def("__repr__", [](const GateCZ& g) { return "CZ(" + std::to_string(g.i) + "," + std::to_string(g.j) + ")"; })
        .def("__str__", [](const GateCZ& g) { return "CZ(" + std::to_string(g.i) + "," + std::to_string(g.j) + ")"; });

    py::class_<RotGate>(m, "ROT")
        .def(py::init([](const std::string& ops, const std::vector<int>& loc, double theta) {
                 if (ops.size() != loc.size()) fail("Invalid arguments (both arguments should have the same size)");
                 std::vector<std::pair<int, char>> s;
                 for (size_t i = 0; i < ops.size(); ++i) s.emplace_back(loc[i], ops[i]);
                 PauliString p;
                 for (auto& f : s) {  // same checks as PauliString(loc, actions) (src/pauli/algebra.h:50-72)
                     if (f.first < 0 || f.first >= MAX_QUBITS)
                         fail("Invalid PauliString (qubit index " + std::to_string(f.first) + " invalid, should be between 0 and " +
                              std::to_string(MAX_QUBITS - 1) + ")");
                     if (f.second != 'I' && f.second != 'X' && f.second != 'Y' && f.second != 'Z')
                         fail(std::string("Invalid PauliString (symbol '") + f.second + "' not recognized)");
                     p.set(f.first, f.second);
                 }
                 return RotGate{p, theta};
             }),
             py::arg("axis"), py::arg("qubits"), py::arg("theta"))
        .def(py::init([](const PauliString& p, double theta) { return RotGate{p, theta}; }))
        .def(py::init([](const std::string& s, double theta) { return RotGate{pauli_from_string(s), theta}; }))
        .def("__call__", [](const RotGate& g, const PauliString& a) { return rot_apply(g, pp_string(a)); })
        .def("__call__", [](const RotGate& g, const PauliPolynomial& p) { return rot_apply(g, p); })
        .def_property_readonly("qubits", [](const RotGate& g) { return g.ps.support(); })
        .def_readonly("axis", &RotGate::ps)
        .def_readonly("theta", &RotGate::theta)
        // cos(theta/2) I - i sin(theta/2) P (src/pauli/gates.h:227-232)
        .def("aspoly", [](const RotGate& g) {
            double c, s;
            trig(g.theta / 2, c, s);
            PauliPolynomial p;
            p.t.at(PauliString()) = c;
            p.t.at(g.ps) += cplx(0, -s);
            return p;
        })
        .def("__repr__", [](const RotGate& g) { return "ROT(" + g.ps.compact() + "," + std::to_string(g.theta) + ")"; })
        .def("__str__", [](const RotGate& g) { return "ROT(" + g.ps.compact() + "," + std::to_string(g.theta) + ")"; });

    // MROT: e^{-i theta/2 i^r m_S}, r = 1 iff deg(S) mod 4 in {2, 3} (src/majorana_gates.h:17-65)
    auto mrot_from_monomial = [](const MajoranaString& s, cplx coeff, double t) {
        const bool sh = s.is_hermitian();
        const bool herm = (sh && coeff.imag() == 0) || (!sh && coeff.real() == 0);
        if (!herm) fail("Supplied Majorana monomial " + fmt_complex(coeff) + "*(" + s.compact() + ") is not Hermitian");
        const bool r = !sh;
        return MRotGate{s, r ? coeff.imag() * t : coeff.real() * t, r};
    };
    py::class_<MRotGate>(m, "MROT")
        .def(py::init([mrot_from_monomial](const std::vector<int>& supp, cplx c, double t) {
            return mrot_from_monomial(majorana_from_support(supp), c, t);
        }))
        .def(py::init([](const MajoranaString& s, double theta) { return MRotGate{s, theta, !s.is_hermitian()}; }))
        .def(py::init([mrot_from_monomial](const MajoranaString& s, cplx c, double t) { return mrot_from_monomial(s, c, t); }))
        .def("__call__", [](const MRotGate& g, const MajoranaString& a) { return mrot_apply(g, mp_string(a)); })
        .def("__call__", [](const MRotGate& g, const MajoranaPolynomial& p) { return mrot_apply(g, p); })
        .def_property_readonly("qubits", [](const MRotGate& g) { return g.ms.support(); })
        .def_readonly("axis", &MRotGate::ms)
        .def_readonly("theta", &MRotGate::theta)
        .def("aspoly", [](const MRotGate& g) {
            double c, s;
            trig(g.theta / 2, c, s);
            MajoranaPolynomial p;
            p.t.at(MajoranaString()) = c;
            p.t.at(g.ms) += g.r ? cplx(s, 0) : cplx(0, -s);
            return p;
        })
        .def("__repr__", [](const MRotGate& g) {
            return std::string("MROT(") + (g.r ? "1j " : "") + g.ms.compact() + ", theta=" + std::to_string(g.theta) + ")";
        })
        .def("__str__", [](const MRotGate& g) {
            return std::string("MROT(") + (g.r ? "1j " : "") + g.ms.compact() + ", theta=" + std::to_string(g.theta) + ")";
        });

    // ---------------- Jordan-Wigner bridge (src/python_bind.cpp:170-216, Majorana/Pauli part) ----------------
    m.def("jw", [](const MajoranaString& a) { return jw_poly(mp_string(a)); },
          "Majorana string to Pauli polynomial (Jordan-Wigner)");
    m.def("jw", [](const MajoranaPolynomial& p) { return jw_poly(p); },
          "Majorana polynomial to Pauli polynomial (Jordan-Wigner)");
    m.def("jw", [](const MRotGate& g) { return jw_rotation(g); },
          "Majorana rotation to the equal Pauli rotation (Jordan-Wigner)");
    m.def("topauli", [](const MajoranaString& a) { return jw_poly(mp_string(a)); });
    m.def("topauli", [](const MajoranaPolynomial& p) { return jw_poly(p); });
    m.def("tomajorana", [](const PauliString& a) { return majorana_poly(pp_string(a)); },
          "Pauli string to Majorana polynomial (inverse Jordan-Wigner)");
    m.def("tomajorana", [](const PauliPolynomial& p) { return majorana_poly(p); },
          "Pauli polynomial to Majorana polynomial (inverse Jordan-Wigner)");
    ms.def("topauli", [](const MajoranaString& a) { return jw_poly(mp_string(a)); });
    mpc.def("topauli", [](const MajoranaPolynomial& p) { return jw_poly(p); });
    ps.def("tomajorana", [](const PauliString& a) { return majorana_poly(pp_string(a)); });
    pp.def("tomajorana", [](const PauliPolynomial& p) { return majorana_poly(p); });

    // ---------------- propagate (two overloads, Pauli first) ----------------
    m.def("propagate", &propagate_pauli, py::arg("circuit"), py::arg("observable"), py::arg("n_threads") = 1,
          py::arg("maxdegree") = py::none(), py::arg("mincoeff") = py::none(), py::arg("topk") = 0,
          py::arg("max_xweight") = -1, py::arg("xtrunc_period") = 1, py::arg("batched") = true,
          py::arg("parallel") = "auto", py::kw_only(), py::arg("device") = 0, py::arg("precision") = "fp64",
          R"DOC(Pauli propagation (Heisenberg picture) on the GPU.

Arguments and defaults are those of fastfermion.propagate; `parallel` and
`n_threads` select CPU strategies in fastfermion and are accepted but unused
here (every strategy gives the same result).  Additive keyword-only arguments:
device (HIP device index), precision ('fp64', 'fp32', 'fp32c').)DOC");
    m.def("propagate", &propagate_majorana, py::arg("circuit"), py::arg("observable"),
          py::arg("maxdegree") = py::none(), py::arg("mincoeff") = 0.0, py::kw_only(), py::arg("device") = 0,
          py::arg("precision") = "fp64", "Majorana propagation on the GPU (fastfermion.propagate overload #2).");
    m.def("propagate_expectation", &propagate_expectation_pauli, py::arg("circuit"), py::arg("observable"),
          py::arg("up") = std::vector<int>{}, py::arg("maxdegree") = py::none(), py::arg("mincoeff") = py::none(),
          py::arg("max_xweight") = -1, py::arg("xtrunc_period") = 1, py::arg("batched") = true, py::kw_only(),
          py::arg("device") = 0, py::arg("precision") = "fp64",
          "<b|U^dag O U|b> for the product state with qubits `up` set, computed on the device without "
          "downloading the propagated terms.");

    // ---------------- profiling (src/python_bind.cpp:427-446) ----------------
    m.def("profile_reset", []() {
        g_last_profile = ffx_profile{};
        g_profile_on = true;
    });
    m.def("profile_disable", []() { g_profile_on = false; });
    m.def("profile_get", []() {
        py::dict d;
        const ffx_profile& p = g_last_profile;
        d["n_gates"] = p.n_gates;
        d["updates"] = p.updates;
        d["peak_terms"] = p.peak_terms;
        d["outs"] = p.outs;
        d["compactions"] = p.compactions;
        d["launches"] = p.launches;
        d["upload"] = p.seconds_upload;
        d["propagate"] = p.seconds_propagate;
        d["download"] = p.seconds_download;
        d["total"] = p.seconds_total;
        // the reference's keys (src/python_bind.cpp:433-446): its CPU phases map to
        // the device ones; emission and merging are one kernel here, and a
        // single device routes every partner locally
        d["snapshot"] = p.seconds_upload;
        d["emit"] = p.seconds_propagate;
        d["merge"] = 0.0;
        d["emit_per_thread"] = py::list();
        d["merge_per_thread"] = py::list();
        d["n_local"] = 0;
        d["n_remote"] = 0;
        return d;
    });
    m.def("device_count", []() { return ffx_device_count(); });

    // operator bases used by examples (src/gen.h:16-60)
    m.def("paulis", [](int n) {
        std::vector<PauliString> x(n), y(n), z(n);
        for (int i = 0; i < n; ++i) {
            check_qubit(i);
            x[i].set(i, 'X');
            y[i].set(i, 'Y');
            z[i].set(i, 'Z');
        }
        return py::make_tuple(x, y, z);
    });
    m.def("majoranas", [](int n) {
        std::vector<MajoranaString> r(n);
        for (int i = 0; i < n; ++i) r[i] = majorana_from_support({i});
        return r;
    });

    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const std::bad_alloc& e) {
            PyErr_SetString(PyExc_MemoryError, "fastfermion_amd: out of memory");
        }
    });
}
