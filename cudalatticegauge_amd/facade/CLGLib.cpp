// CLGLib.cpp -- implementation of the CLGLib-named facade over libclg_hip.so (see CLGLib.h).
#include "CLGLib.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace CLGLib {

namespace {

CLatticeData* g_lattice = nullptr;

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) ++a;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
    return s.substr(a, b - a);
}

void check(int st, const char* what) {
    if (st != CLG_OK && st != CLG_NOT_CONVERGED) {
        std::fprintf(stderr, "[CLGLib] %s failed: %s: %s\n", what, clg_status_string(st), clg_last_error());
        throw std::runtime_error(what);
    }
}

std::vector<std::string> split_array(const std::string& v) {
    std::vector<std::string> out;
    std::string t = trim(v);
    if (t.size() < 2 || t.front() != '[' || t.back() != ']') return out;
    t = t.substr(1, t.size() - 2);
    std::stringstream ss(t);
    std::string item;
    while (std::getline(ss, item, ',')) {
        item = trim(item);
        if (!item.empty()) out.push_back(item);
    }
    return out;
}

}  // namespace

// ================================================================ parameters
UBOOL CParameters::Exist(const std::string& key) const { return m_values.count(key) || m_children.count(key); }
UBOOL CParameters::FetchStringValue(const std::string& key, std::string& v) const {
    auto it = m_values.find(key);
    if (it == m_values.end()) return false;
    v = it->second;
    return true;
}
UBOOL CParameters::FetchValueINT(const std::string& key, INT& v) const {
    std::string s;
    if (!FetchStringValue(key, s)) return false;
    try {
        v = std::stoi(s);
    } catch (...) {
        return false;
    }
    return true;
}
UBOOL CParameters::FetchValueReal(const std::string& key, Real& v) const {
    std::string s;
    if (!FetchStringValue(key, s)) return false;
    try {
        v = std::stod(s);
    } catch (...) {
        return false;
    }
    return true;
}
UBOOL CParameters::FetchValueArrayReal(const std::string& key, std::vector<Real>& v) const {
    std::string s;
    if (!FetchStringValue(key, s)) return false;
    v.clear();
    for (const std::string& x : split_array(s)) v.push_back(std::stod(x));
    return true;
}
UBOOL CParameters::FetchValueArrayINT(const std::string& key, std::vector<INT>& v) const {
    std::string s;
    if (!FetchStringValue(key, s)) return false;
    v.clear();
    for (const std::string& x : split_array(s)) v.push_back(std::stoi(x));
    return true;
}
UBOOL CParameters::FetchParameterValue(const std::string& key, CParameters& sub) const {
    auto it = m_children.find(key);
    if (it == m_children.end()) return false;
    sub = it->second;
    return true;
}
std::vector<std::string> CParameters::ChildNames() const {
    std::vector<std::string> out;
    for (auto& kv : m_children) out.push_back(kv.first);
    return out;
}

UBOOL CYAMLParser::ParseText(const std::string& text, CParameters& out) {
    struct Level {
        int indent;
        CParameters* node;
    };
    std::vector<Level> stack{{-1, &out}};
    std::stringstream ss(text);
    std::string line;
    std::string pending_key;
    int pending_indent = -1;
    while (std::getline(ss, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        int indent = 0;
        while (indent < (int)line.size() && (line[indent] == ' ' || line[indent] == '\t'))
            indent += line[indent] == '\t' ? 4 : 1;
        const std::string body = trim(line);
        if (body.empty()) continue;
        const size_t colon = body.find(':');
        if (colon == std::string::npos) continue;  // not a mapping line
        const std::string key = trim(body.substr(0, colon));
        const std::string val = trim(body.substr(colon + 1));
        if (!pending_key.empty()) {
            // a key without value opens a map when the next line is indented deeper
            if (indent > pending_indent) {
                CParameters* child = &stack.back().node->Child(pending_key);
                stack.push_back({pending_indent, child});
            } else {
                stack.back().node->SetStringValue(pending_key, "");
            }
            pending_key.clear();
        }
        while (stack.size() > 1 && stack.back().indent >= indent) stack.pop_back();
        if (val.empty()) {
            pending_key = key;
            pending_indent = indent;
        } else {
            stack.back().node->SetStringValue(key, val);
        }
    }
    if (!pending_key.empty()) stack.back().node->SetStringValue(pending_key, "");
    return true;
}
UBOOL CYAMLParser::ParseFile(const std::string& path, CParameters& out) {
    std::ifstream f(path);
    if (!f) return false;
    std::stringstream buf;
    buf << f.rdbuf();
    return ParseText(buf.str(), out);
}

// ================================================================ fields
CField::CField(CLatticeData* owner, int kind) : m_pOwner(owner), m_kind(kind) {
    check(clg_field_create(owner->m_ctx, kind, CLG_F64, CLG_FULL, owner->m_dims, &m_h), "field create");
}
CField::~CField() {
    if (m_h) clg_field_destroy(m_h);
}
void CField::Zero() { check(clg_field_zero(m_h), "Zero"); }
void CField::CopyTo(CField* dst) const { check(clg_field_copy(dst->m_h, m_h), "CopyTo"); }
void CField::Axpy(const CLGComplex& a, const CField* x) {
    if (m_kind == CLG_GAUGE) {
        check(clg_gauge_axpy(m_pOwner->m_ctx, m_h, a.x, x->m_h), "Axpy");
        return;
    }
    check(clg_axpy(m_pOwner->m_ctx, m_h, a.x, a.y, x->m_h), "Axpy");
}
void CField::Axpy(Real a, const CField* x) { Axpy(_make_cuComplex(a, 0.0), x); }
void CField::AxpyPlus(const CField* x) { Axpy(1.0, x); }
void CField::AxpyMinus(const CField* x) { Axpy(-1.0, x); }
void CField::ScalarMultply(const CLGComplex& a) { check(clg_scale(m_pOwner->m_ctx, m_h, a.x, a.y), "Scale"); }
CLGComplex CField::Dot(const CField* other) const {
    double z[2];
    check(clg_dot(m_pOwner->m_ctx, m_h, other->m_h, z), "Dot");
    return _make_cuComplex(z[0], z[1]);
}
std::vector<double> CField::Download() const {
    std::vector<double> v((size_t)clg_field_host_doubles(m_h));
    check(clg_field_download(m_h, v.data()), "download");
    return v;
}
void CField::Upload(const std::vector<double>& host) { check(clg_field_upload(m_h, host.data()), "upload"); }

CFieldGaugeSU3::CFieldGaugeSU3(CLatticeData* owner) : CFieldGauge(owner, CLG_GAUGE) {}
void CFieldGaugeSU3::InitialField(EFieldInitialType t, unsigned seed, Real eps) {
    if (t == EFIT_Identity) check(clg_gauge_unit(m_h), "gauge unit");
    else if (t == EFIT_Zero) check(clg_field_zero(m_h), "gauge zero");
    else check(clg_gauge_random(m_h, eps, seed), "gauge random");
}
bool appFileTypeFromString(const std::string& name, EFieldFileType& out) {
    static const std::pair<const char*, EFieldFileType> names[] = {
        {"EFFT_BridgePPTXT", EFFT_BridgePPTXT}, {"EFFT_BridgePPBin", EFFT_BridgePPBin},
        {"EFFT_CLGBin", EFFT_CLGBin},           {"EFFT_CLGBinCompressed", EFFT_CLGBinCompressed},
        {"EFFT_CLGBinFloat", EFFT_CLGBinFloat}, {"EFFT_CLGBinDouble", EFFT_CLGBinDouble}};
    for (const auto& n : names)
        if (name == n.first) {
            out = n.second;
            return true;
        }
    return false;
}
void CField::InitialFieldWithFile(const std::string& fileName, EFieldFileType type) {
    check(clg_field_load_file(m_h, fileName.c_str(), (int)type), "InitialFieldWithFile");
}
std::string CField::SaveToFile(const std::string& fileName, EFieldFileType type) const {
    char md5[33] = {0};
    check(clg_field_save_file(m_h, fileName.c_str(), (int)type, md5), "SaveToFile");
    return md5;
}

DOUBLE CFieldGaugeSU3::CalculatePlaqutteEnergy(DOUBLE betaOverN) const {
    double e = 0;
    check(clg_gauge_plaquette_energy(m_pOwner->m_ctx, m_h, 3.0 * betaOverN, &e), "plaquette");
    return e;
}
CField* CFieldGaugeSU3::GetCopy() const {
    CFieldGaugeSU3* c = new CFieldGaugeSU3(m_pOwner);
    CopyTo(c);
    return c;
}
std::string CFieldGaugeSU3::SaveToCompressedFile(const std::string& fileName) const {
    return SaveToFile(fileName, EFFT_CLGBinCompressed);
}

CFieldFermion::CFieldFermion(CLatticeData* owner, int kind, Real param, const int bc[4])
    : CField(owner, kind), m_fParam(param) {
    for (int i = 0; i < 4; ++i) m_bc[i] = bc[i];
    check(clg_field_set_bc(m_h, bc), "set bc");
}
const CFieldGaugeSU3* CFieldFermion::GetDefaultGauge(INT gaugeNum, const CFieldGauge* const* gaugeFields) const {
    for (INT i = 0; i < gaugeNum; ++i)
        if (gaugeFields[i] && gaugeFields[i]->m_byFieldId == (int)m_byGaugeFieldId)
            return dynamic_cast<const CFieldGaugeSU3*>(gaugeFields[i]);
    return nullptr;
}
namespace {
const CFieldGaugeSU3* need_gauge(const CFieldFermion* f, INT gaugeNum, const CFieldGauge* const* gaugeFields) {
    const CFieldGaugeSU3* g = f->GetDefaultGauge(gaugeNum, gaugeFields);
    if (!g) throw std::runtime_error("no SU(3) gauge field with the fermion's gauge field id among the gauge fields");
    return g;
}
}  // namespace
// this <- coeff * op(this), op one of D, D^dagger, DD, DD^dagger, with the given mass / hopping
void CFieldFermion::Apply(EFieldOperator op, const CFieldGaugeSU3* gauge, Real param, EOperatorCoefficientType ct,
                          Real re, Real im) {
    std::unique_ptr<CField> tmp(GetCopy());
    if (Kind() == CLG_WILSON)
        check(clg_wilson_apply(m_pOwner->m_ctx, m_h, tmp->Handle(), gauge->Handle(), param, op, ct, re, im), "D");
    else
        check(clg_ks_apply(m_pOwner->m_ctx, m_h, tmp->Handle(), gauge->Handle(), param, op, ct, re, im), "D");
}
void CFieldFermion::D(INT gn, INT bn, const CFieldGauge* const* gf, const CFieldBoson* const*,
                      EOperatorCoefficientType t, Real re, Real im) {
    Apply(EFO_F_D, need_gauge(this, gn, gf), m_fParam, t, re, im);
    (void)bn;
}
void CFieldFermion::Ddagger(INT gn, INT, const CFieldGauge* const* gf, const CFieldBoson* const*,
                            EOperatorCoefficientType t, Real re, Real im) {
    Apply(EFO_F_Ddagger, need_gauge(this, gn, gf), m_fParam, t, re, im);
}
void CFieldFermion::DD(INT gn, INT, const CFieldGauge* const* gf, const CFieldBoson* const*, EOperatorCoefficientType t,
                       Real re, Real im) {
    Apply(EFO_F_DD, need_gauge(this, gn, gf), m_fParam, t, re, im);
}
void CFieldFermion::DDdagger(INT gn, INT, const CFieldGauge* const* gf, const CFieldBoson* const*,
                             EOperatorCoefficientType t, Real re, Real im) {
    Apply(EFO_F_DDdagger, need_gauge(this, gn, gf), m_fParam, t, re, im);
}
void CFieldFermion::DWithMass(INT gn, INT, const CFieldGauge* const* gf, const CFieldBoson* const*, Real m,
                              EOperatorCoefficientType t, Real re, Real im) {
    Apply(EFO_F_D, need_gauge(this, gn, gf), m, t, re, im);
}
void CFieldFermion::DdaggerWithMass(INT gn, INT, const CFieldGauge* const* gf, const CFieldBoson* const*, Real m,
                                    EOperatorCoefficientType t, Real re, Real im) {
    Apply(EFO_F_Ddagger, need_gauge(this, gn, gf), m, t, re, im);
}
void CFieldFermion::DDWithMass(INT gn, INT, const CFieldGauge* const* gf, const CFieldBoson* const*, Real m,
                               EOperatorCoefficientType t, Real re, Real im) {
    Apply(EFO_F_DD, need_gauge(this, gn, gf), m, t, re, im);
}
void CFieldFermion::DDdaggerWithMass(INT gn, INT, const CFieldGauge* const* gf, const CFieldBoson* const*, Real m,
                                     EOperatorCoefficientType t, Real re, Real im) {
    Apply(EFO_F_DDdagger, need_gauge(this, gn, gf), m, t, re, im);
}
UBOOL CFieldFermion::ApplyOperator(EFieldOperator op, INT gn, INT bn, const CFieldGauge* const* gf,
                                   const CFieldBoson* const* bf, EOperatorCoefficientType t, Real re, Real im,
                                   void* other) {
    switch ((int)op) {
        case EFO_F_D: D(gn, bn, gf, bf, t, re, im); return true;
        case EFO_F_Ddagger: Ddagger(gn, bn, gf, bf, t, re, im); return true;
        case EFO_F_DD: DD(gn, bn, gf, bf, t, re, im); return true;
        case EFO_F_DDdagger: DDdagger(gn, bn, gf, bf, t, re, im); return true;
        case EFO_F_InverseD:
        case EFO_F_InverseDdagger:
        case EFO_F_InverseDD:
        case EFO_F_InverseDDdagger:
            if (t != EOCT_None) {
                std::fprintf(stderr, "[CLGLib] ApplyOperator: an inverse with a coefficient is not implemented\n");
                return false;
            }
            return Inverse(op, gn, bn, gf, bf);
        case EFO_F_RationalD:
            if (!other) {
                std::fprintf(stderr, "[CLGLib] ApplyOperator: EFO_F_RationalD needs a CRatinalApproximation\n");
                return false;
            }
            return RationalApproximation(EFO_F_D, gn, bn, gf, bf, static_cast<const CRatinalApproximation*>(other));
        case EFO_F_D_WithMass:
        case EFO_F_Ddagger_WithMass:
        case EFO_F_DD_WithMass:
        case EFO_F_DDdagger_WithMass: {
            if (!other) {
                std::fprintf(stderr, "[CLGLib] ApplyOperator: a *_WithMass operator needs the mass\n");
                return false;
            }
            const Real m = *static_cast<const Real*>(other);
            const EFieldOperator base = (EFieldOperator)((int)op - (int)EFO_F_D_WithMass);
            Apply(base, need_gauge(this, gn, gf), m, t, re, im);
            return true;
        }
        default:
            std::fprintf(stderr, "[CLGLib] ApplyOperator: operator %d is not implemented\n", (int)op);
            return false;
    }
}
UBOOL CFieldFermion::Inverse(EFieldOperator op, INT gn, INT bn, const CFieldGauge* const* gf,
                             const CFieldBoson* const* bf) {
    CSLASolver* s = appGetFermionSolver(m_byFieldId);
    if (!s) {
        std::fprintf(stderr, "[CLGLib] no solver for field %d\n", m_byFieldId);
        return false;
    }
    return s->Solve(this, this, gn, bn, gf, bf, op);
}
UBOOL CFieldFermion::InverseD(INT gn, INT bn, const CFieldGauge* const* gf, const CFieldBoson* const* bf) {
    return Inverse(EFO_F_InverseD, gn, bn, gf, bf);
}
UBOOL CFieldFermion::InverseDdagger(INT gn, INT bn, const CFieldGauge* const* gf, const CFieldBoson* const* bf) {
    return Inverse(EFO_F_InverseDdagger, gn, bn, gf, bf);
}
UBOOL CFieldFermion::InverseDD(INT gn, INT bn, const CFieldGauge* const* gf, const CFieldBoson* const* bf) {
    return Inverse(EFO_F_InverseDD, gn, bn, gf, bf);
}
UBOOL CFieldFermion::InverseDDdagger(INT gn, INT bn, const CFieldGauge* const* gf, const CFieldBoson* const* bf) {
    return Inverse(EFO_F_InverseDDdagger, gn, bn, gf, bf);
}
void CFieldFermion::PrepareForHMC(INT gn, INT, const CFieldGauge* const* gf, const CFieldBoson* const*) {
    // the heat-bath draw is seeded per lattice and field id
    PrepareForHMCS(need_gauge(this, gn, gf), m_pOwner->m_uiSeed * 31u + (unsigned)m_byFieldId + 7u * m_uiHeatBath++);
}
UBOOL CFieldFermion::CalculateForce(INT gn, INT bn, const CFieldGauge* const* gf, const CFieldBoson* const*,
                                    CFieldGauge* const* gforce, CFieldBoson* const*, ESolverPhase) const {
    const CFieldGaugeSU3* g = need_gauge(this, gn, gf);
    INT idx = -1;
    for (INT i = 0; i < gn; ++i)
        if (gf[i] == g) idx = i;
    CFieldGaugeSU3* F = idx >= 0 ? dynamic_cast<CFieldGaugeSU3*>(gforce[idx]) : nullptr;
    if (!F) throw std::runtime_error("CalculateForce: no SU(3) force field for the fermion's gauge field");
    clg_solve_report rep{};
    int st;
    if (Kind() == CLG_WILSON) {
        CSLASolver* s = appGetFermionSolver(m_byFieldId);
        if (!s) throw std::runtime_error("CalculateForce: no solver for the field");
        clg_solver_params p = s->m_params;
        p.kind = CLG_WILSON;
        st = clg_wilson_force(m_pOwner->m_ctx, &p, F->Handle(), m_h, g->Handle(), m_fParam, nullptr, nullptr, &rep);
    } else {
        const CFieldFermionKSSU3* ks = static_cast<const CFieldFermionKSSU3*>(this);
        CMultiShiftSolver* ms = appGetMultiShiftSolver(m_byFieldId);
        if (!ms || ks->m_rMD.size() < 3) throw std::runtime_error("CalculateForce: the KS field needs an MSSolver and MD");
        clg_solver_params p = ms->m_params;
        p.kind = CLG_KS;
        st = clg_ks_force(m_pOwner->m_ctx, &p, ms->m_kind, ms->m_prec, F->Handle(), m_h, g->Handle(), m_fParam,
                          ks->m_rMD.data(), (int)(ks->m_rMD.size() - 1) / 2, &rep);
    }
    check(st, "CalculateForce");
    (void)bn;
    return st == CLG_OK;
}
void CFieldFermion::InitialField(EFieldInitialType t, unsigned seed) {
    if (t == EFIT_RandomGaussian) check(clg_field_gaussian(m_h, seed), "gaussian");
    else check(clg_field_zero(m_h), "zero");
}
void CFieldFermion::InitialField(EFieldInitialType t) { InitialField(t, m_pOwner->NextSeed()); }

namespace {
// columns of the dense operator: apply(out, in) to every point source of the field's lattice,
// downloaded as complex columns of n = nsites * ncomp entries
template <class Apply>
std::vector<std::vector<CLGComplex>> dense_columns(clg_field in, clg_field out, const int dims[4], int nspin,
                                                   Apply apply) {
    const long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
    const int nc = 3 * nspin;
    const long n = V * nc;
    std::vector<std::vector<CLGComplex>> cols((size_t)n);
    std::vector<double> host((size_t)(2 * n));
    for (long site = 0; site < V; ++site) {
        int c[4];
        long r = site;
        for (int d = 3; d >= 0; --d) {
            c[d] = (int)(r % dims[d]);
            r /= dims[d];
        }
        for (int k = 0; k < nc; ++k) {
            check(clg_field_zero(in), "zero");
            check(clg_field_point_source(in, c, nspin == 4 ? k / 3 : 0, k % 3), "point source");
            apply(out, in);
            check(clg_field_download(out, host.data()), "download");
            std::vector<CLGComplex>& col = cols[(size_t)(site * nc + k)];
            col.resize((size_t)n);
            for (long i = 0; i < n; ++i) col[(size_t)i] = _make_cuComplex(host[(size_t)(2 * i)], host[(size_t)(2 * i + 1)]);
        }
    }
    return cols;
}
double cabs(const CLGComplex& z) { return std::sqrt(z.x * z.x + z.y * z.y); }
}  // namespace

UINT CFieldFermionWilsonSquareSU3::TestGamma5Hermitian(const CFieldGauge* pGauge, UBOOL bTestGamma5) {
    const CFieldGaugeSU3* g = dynamic_cast<const CFieldGaugeSU3*>(pGauge);
    if (!g) throw std::runtime_error("TestGamma5Hermitian needs an SU(3) gauge field");
    clg_context ctx = m_pOwner->m_ctx;
    const int* dims = m_pOwner->m_dims;
    clg_field in = nullptr, out = nullptr;
    check(clg_field_create(ctx, CLG_WILSON, CLG_F64, CLG_FULL, dims, &in), "field");
    check(clg_field_create(ctx, CLG_WILSON, CLG_F64, CLG_FULL, dims, &out), "field");
    clg_field_set_bc(in, m_bc);
    auto applier = [&](int op) {
        return [&, op](clg_field o, clg_field i) {
            check(clg_wilson_apply(ctx, o, i, g->Handle(), m_fParam, op, EOCT_None, 1.0, 0.0), "D");
        };
    };
    const auto Dm = dense_columns(in, out, dims, 4, applier(EFO_F_D));
    const long n = (long)Dm.size();
    double big = 0, worst = 0;
    for (const auto& col : Dm)
        for (const auto& z : col) big = std::max(big, cabs(z));
    if (bTestGamma5) {
        // the implemented D^dagger against the conjugate transpose of D
        const auto Dd = dense_columns(in, out, dims, 4, applier(EFO_F_Ddagger));
        for (long j = 0; j < n; ++j)
            for (long i = 0; i < n; ++i) {
                const CLGComplex a = Dd[(size_t)j][(size_t)i], b = Dm[(size_t)i][(size_t)j];
                worst = std::max(worst, std::hypot(a.x - b.x, a.y + b.y));
            }
    } else {
        // gamma5 D Hermitian: g5_i D_ij = conj(g5_j D_ji), g5 = +1 on spins 0,1 and -1 on spins 2,3
        auto g5 = [](long idx) { return (idx % 12) < 6 ? 1.0 : -1.0; };
        for (long j = 0; j < n; ++j)
            for (long i = 0; i < n; ++i) {
                const CLGComplex a = Dm[(size_t)j][(size_t)i], b = Dm[(size_t)i][(size_t)j];
                const double si = g5(i), sj = g5(j);
                worst = std::max(worst, std::hypot(si * a.x - sj * b.x, si * a.y + sj * b.y));
            }
    }
    clg_field_destroy(in);
    clg_field_destroy(out);
    m_fLastTestError = big > 0 ? worst / big : worst;
    return m_fLastTestError > 1e-12 ? 1u : 0u;
}

UINT CFieldFermionKSSU3::TestAntiHermitian(INT gn, INT, const CFieldGauge* const* gf, const CFieldBoson* const*) {
    const CFieldGaugeSU3* g = need_gauge(this, gn, gf);
    clg_context ctx = m_pOwner->m_ctx;
    const int* dims = m_pOwner->m_dims;
    clg_field in = nullptr, out = nullptr;
    check(clg_field_create(ctx, CLG_KS, CLG_F64, CLG_FULL, dims, &in), "field");
    check(clg_field_create(ctx, CLG_KS, CLG_F64, CLG_FULL, dims, &out), "field");
    clg_field_set_bc(in, m_bc);
    const auto D0 = dense_columns(in, out, dims, 1, [&](clg_field o, clg_field i) {
        check(clg_ks_apply(ctx, o, i, g->Handle(), 0.0, EFO_F_D, EOCT_None, 1.0, 0.0), "D");
    });
    const long n = (long)D0.size();
    double big = 0, worst = 0;
    for (const auto& col : D0)
        for (const auto& z : col) big = std::max(big, cabs(z));
    for (long j = 0; j < n; ++j)
        for (long i = 0; i < n; ++i) {
            const CLGComplex a = D0[(size_t)j][(size_t)i], b = D0[(size_t)i][(size_t)j];
            worst = std::max(worst, std::hypot(a.x + b.x, a.y - b.y));  // D_ij + conj(D_ji)
        }
    clg_field_destroy(in);
    clg_field_destroy(out);
    m_fLastTestError = big > 0 ? worst / big : worst;
    return m_fLastTestError > 1e-12 ? 1u : 0u;
}

CGaugeFixingRandom::CGaugeFixingRandom(CLatticeData* owner, Real eps) {
    check(clg_field_create(owner->m_ctx, CLG_GAUGE, CLG_F64, CLG_FULL, owner->m_dims, &m_gt), "field");
    check(clg_gauge_transform_random(m_gt, eps, owner->NextSeed()), "gauge transform");
}
CGaugeFixingRandom::~CGaugeFixingRandom() {
    if (m_gt) clg_field_destroy(m_gt);
}
void CGaugeFixingRandom::GaugeFixing(CFieldGauge* pGauge) {
    check(clg_gauge_transform_apply(m_gt, pGauge->Handle()), "GaugeFixing");
}
void CGaugeFixingRandom::AlsoFixingFermion(CFieldFermion* pFermion) {
    check(clg_gauge_transform_apply(m_gt, pFermion->Handle()), "AlsoFixingFermion");
}
void CFieldFermion::InitialAsSource(const int coord[4], int spin, int colour) {
    check(clg_field_point_source(m_h, coord, spin, colour), "point source");
}

CFieldFermionWilsonSquareSU3::CFieldFermionWilsonSquareSU3(CLatticeData* owner, Real kappa, const int bc[4])
    : CFieldFermion(owner, CLG_WILSON, kappa, bc) {}
void CFieldFermionWilsonSquareSU3::PrepareForHMCS(const CFieldGaugeSU3* gauge, unsigned seed) {
    check(clg_wilson_prepare_hmc(m_pOwner->m_ctx, m_h, gauge->Handle(), m_fParam, seed), "PrepareForHMC");
}
CField* CFieldFermionWilsonSquareSU3::GetCopy() const {
    CFieldFermionWilsonSquareSU3* c = new CFieldFermionWilsonSquareSU3(m_pOwner, m_fParam, m_bc);
    c->m_byFieldId = m_byFieldId;
    CopyTo(c);
    return c;
}

CFieldFermionKSSU3::CFieldFermionKSSU3(CLatticeData* owner, Real mass, const int bc[4])
    : CFieldFermion(owner, CLG_KS, mass, bc) {}
void CFieldFermionKSSU3::PrepareForHMCS(const CFieldGaugeSU3* gauge, unsigned seed) {
    CMultiShiftSolver* ms = appGetMultiShiftSolver(m_byFieldId);
    if (!ms || m_rMC.size() < 3) throw std::runtime_error("KS PrepareForHMC needs an MSSolver and MC");
    clg_solve_report rep{};
    check(clg_ks_prepare_hmc(m_pOwner->m_ctx, &ms->m_params, ms->m_kind, ms->m_prec, m_h, gauge->Handle(), m_fParam,
                             m_rMC.data(), (int)(m_rMC.size() - 1) / 2, seed, &rep),
          "KS PrepareForHMC");
}
UBOOL CFieldFermion::RationalApproximation(EFieldOperator op, INT gn, INT, const CFieldGauge* const* gf,
                                           const CFieldBoson* const*, const CRatinalApproximation* r) {
    const CFieldGaugeSU3* gauge = need_gauge(this, gn, gf);
    CMultiShiftSolver* ms = appGetMultiShiftSolver(m_byFieldId);
    if (!ms || !r || r->m_c.size() < 3 || r->m_c.size() % 2 == 0) {
        std::fprintf(stderr, "[CLGLib] RationalApproximation needs an MSSolver and [c, a_1..a_n, b_1..b_n]\n");
        return false;
    }
    clg_solve_report rep{};
    clg_solver_params p = ms->m_params;
    p.kind = Kind();
    const int st = clg_rational_apply(m_pOwner->m_ctx, &p, ms->m_kind, ms->m_prec, m_h, m_h, gauge->Handle(), m_fParam,
                                      (int)op, r->m_c.data(), r->Degree(), &rep);
    check(st, "RationalApproximation");
    return st == CLG_OK;
}
void CFieldFermionKSSU3::D_MC(INT gn, INT bn, const CFieldGauge* const* gf, const CFieldBoson* const* bf) {
    const CRatinalApproximation r(m_rMC);
    RationalApproximation(EFO_F_DDdagger, gn, bn, gf, bf, &r);
}
void CFieldFermionKSSU3::D_MD(INT gn, INT bn, const CFieldGauge* const* gf, const CFieldBoson* const* bf) {
    const CRatinalApproximation r(m_rMD);
    RationalApproximation(EFO_F_DDdagger, gn, bn, gf, bf, &r);
}
CField* CFieldFermionKSSU3::GetCopy() const {
    CFieldFermionKSSU3* c = new CFieldFermionKSSU3(m_pOwner, m_fParam, m_bc);
    c->m_byFieldId = m_byFieldId;
    c->m_rMC = m_rMC;
    c->m_rMD = m_rMD;
    CopyTo(c);
    return c;
}

// ================================================================ solvers
// YAML keys of CSLASolverGMRES / CSLASolverBiCGStab::Configurate (CSolverGMRES.cpp:35-99,
// CSolverBiCGstab.cpp:33-93); EvenOdd, InnerPrecision and InnerAccuracy are this library's keys
void CSLASolver::Configurate(const CParameters& p) {
    INT iv;
    Real fv;
    if (p.FetchValueINT("MaxDim", iv)) m_params.max_dim = iv;
    if (p.FetchValueINT("Restart", iv)) m_params.restart = iv;
    if (p.FetchValueINT("MaxStep", iv)) m_params.max_step = iv;
    if (p.FetchValueINT("DiviationStep", iv)) m_params.deviation_step = iv;
    if (p.FetchValueReal("Accuracy", fv)) m_params.accuracy = fv;
    if (p.FetchValueINT("AbsoluteAccuracy", iv)) m_params.absolute_accuracy = iv;
    if (p.FetchValueINT("EvenOdd", iv)) m_params.even_odd = iv;
    std::string s;
    if (p.FetchStringValue("InnerPrecision", s)) m_params.inner_prec = s == "FP32" ? CLG_F32 : s == "FP16" ? CLG_F16 : CLG_F64;
    if (p.FetchValueReal("InnerAccuracy", fv)) m_params.inner_accuracy = fv;
}
UBOOL CSLASolver::Solve(CField* x, const CField* b, INT gn, INT, const CFieldGauge* const* gf,
                        const CFieldBoson* const*, EFieldOperator op, ESolverPhase, const CField* start) {
    const CFieldFermion* fb = dynamic_cast<const CFieldFermion*>(b);
    if (!fb) throw std::runtime_error("solver needs a fermion field");
    const CFieldGaugeSU3* gauge = need_gauge(fb, gn, gf);
    clg_solver_params p = m_params;
    p.kind = dynamic_cast<const CFieldFermionKSSU3*>(b) ? CLG_KS : CLG_WILSON;
    if (p.kind == CLG_KS) p.even_odd = 0;
    const int iop = op < EFO_F_InverseD ? op + 4 : op;
    const int st = clg_solve(x->Lattice()->m_ctx, &p, x->Handle(), b->Handle(), gauge->Handle(), fb->Param(), iop,
                             start ? start->Handle() : nullptr, &m_lastReport);
    check(st, "Solve");
    return st == CLG_OK;
}
CSLASolverGMRES::CSLASolverGMRES() {
    clg_solver_params_default(&m_params);
    m_params.solver = CLG_SOLVER_GMRES;
    m_params.even_odd = 1;
}
CSLASolverBiCGStab::CSLASolverBiCGStab() {
    clg_solver_params_default(&m_params);
    m_params.solver = CLG_SOLVER_BICGSTAB;
}
CSLASolverGCRODR::CSLASolverGCRODR() { clg_gcrodr_params_default(&m_gparams); }
CSLASolverGCRODR::~CSLASolverGCRODR() {
    if (m_solver) clg_gcrodr_destroy(m_solver);
}
// CSLASolverGCRODR::Configurate keys (CSolverGCRODR.cpp:67-118)
void CSLASolverGCRODR::Configurate(const CParameters& p) {
    INT iv;
    Real fv;
    if (p.FetchValueINT("MDim", iv)) m_gparams.m_dim = iv;
    if (p.FetchValueINT("KDim", iv)) m_gparams.k_dim = iv;
    else m_gparams.k_dim = m_gparams.m_dim - 2;
    if (p.FetchValueINT("Restart", iv)) m_gparams.restart = iv;
    if (p.FetchValueINT("RecalculateR", iv)) m_gparams.recalculate_r = iv;
    if (p.FetchValueINT("AbsoluteAccuracy", iv)) m_gparams.absolute_accuracy = iv;
    if (p.FetchValueReal("Accuracy", fv)) m_gparams.accuracy = fv;
    std::string s;
    if (p.FetchStringValue("DeflateSpace", s))
        m_gparams.deflate = s == "EEDT_REV" ? CLG_DEFLATE_REV : s == "EEDT_HEV" ? CLG_DEFLATE_HEV : CLG_DEFLATE_SVD;
}
UBOOL CSLASolverGCRODR::Solve(CField* x, const CField* b, INT gn, INT, const CFieldGauge* const* gf,
                              const CFieldBoson* const*, EFieldOperator op, ESolverPhase phase, const CField* start) {
    const CFieldFermion* fb = dynamic_cast<const CFieldFermion*>(b);
    if (!fb) throw std::runtime_error("solver needs a fermion field");
    const CFieldGaugeSU3* gauge = need_gauge(fb, gn, gf);
    const int kind = dynamic_cast<const CFieldFermionKSSU3*>(b) ? CLG_KS : CLG_WILSON;
    if (!m_solver || kind != m_gparams.kind || m_ctx != x->Lattice()->m_ctx) {
        if (m_solver) clg_gcrodr_destroy(m_solver);
        m_gparams.kind = kind;
        m_ctx = x->Lattice()->m_ctx;
        check(clg_gcrodr_create(m_ctx, &m_gparams, &m_solver), "GCRODR create");
    }
    const int fop = op >= EFO_F_InverseD ? op - 4 : op;
    // x may alias b: solve into a copy
    std::unique_ptr<CField> xt(b->GetCopy());
    const int st = clg_gcrodr_solve(m_solver, xt->Handle(), b->Handle(), gauge->Handle(), fb->Param(), fop, phase,
                                    start ? start->Handle() : nullptr, &m_lastReport);
    check(st, "GCRODR Solve");
    xt->CopyTo(x);
    return st == CLG_OK;
}

CMultiShiftSolver::CMultiShiftSolver(int kind) : m_kind(kind) {
    clg_solver_params_default(&m_params);
    m_params.max_dim = 30;
    m_params.restart = 100;
}
void CMultiShiftSolver::Configurate(const CParameters& p) {
    INT iv;
    Real fv;
    if (p.FetchValueINT("MaxDim", iv)) m_params.max_dim = iv;
    if (p.FetchValueINT("Restart", iv)) m_params.restart = iv;
    if (p.FetchValueINT("MaxStep", iv)) m_params.max_step = iv;
    if (p.FetchValueINT("DiviationStep", iv)) m_params.deviation_step = iv;
    if (p.FetchValueReal("Accuracy", fv)) m_params.accuracy = fv;
    if (p.FetchValueINT("AbsoluteAccuracy", iv)) m_params.absolute_accuracy = iv;
    std::string s;
    if (p.FetchStringValue("Precision", s)) m_prec = s == "FP32" ? CLG_F32 : CLG_F64;
    // the nested solver of CMultiShiftNested
    CParameters nested;
    if (p.FetchParameterValue("NestedSolver", nested)) {
        CSLASolverGMRES g;
        g.Configurate(nested);
        m_params.max_dim = g.m_params.max_dim;
        m_params.restart = g.m_params.restart;
        m_params.accuracy = g.m_params.accuracy;
        m_params.absolute_accuracy = g.m_params.absolute_accuracy;
    }
}
UBOOL CMultiShiftSolver::Solve(TArray<CField*>& X, const TArray<CLGComplex>& cn, const CField* b, INT gn, INT,
                               const CFieldGauge* const* gf, const CFieldBoson* const*, EFieldOperator op,
                               ESolverPhase, const CField*) {
    const CFieldFermion* fb = dynamic_cast<const CFieldFermion*>(b);
    if (!fb) throw std::runtime_error("solver needs a fermion field");
    if (X.Num() != cn.Num()) throw std::runtime_error("one solution field per shift");
    const CFieldGaugeSU3* gauge = need_gauge(fb, gn, gf);
    clg_solver_params p = m_params;
    p.kind = dynamic_cast<const CFieldFermionKSSU3*>(b) ? CLG_KS : CLG_WILSON;
    std::vector<clg_field> xs;
    std::vector<double> sh;
    for (INT i = 0; i < X.Num(); ++i) {
        xs.push_back(X[i]->Handle());
        sh.push_back(cn[i].x);
        sh.push_back(cn[i].y);
    }
    clg_solve_report rep{};
    const int st = clg_multishift_solve(b->Lattice()->m_ctx, &p, m_kind, m_prec, xs.data(), sh.data(), X.Num(),
                                        b->Handle(), gauge->Handle(), fb->Param(), op, &rep);
    check(st, "MultiShift Solve");
    return st == CLG_OK;
}

// ================================================================ actions
void CActionGaugePlaquette::Initialize(CLatticeData* owner, const CParameters& p) {
    m_pOwner = owner;
    Real beta = 5.0;
    p.FetchValueReal("Beta", beta);
    m_fBetaOverN = beta / 3.0;
}
clg_action CActionGaugePlaquette::Describe() const {
    clg_action a;
    std::memset(&a, 0, sizeof(a));
    a.kind = CLG_ACTION_GAUGE_PLAQUETTE;
    a.beta = 3.0 * m_fBetaOverN;
    return a;
}
DOUBLE CActionGaugePlaquette::EnergySingleField(const CFieldGaugeSU3* gauge) const {
    return gauge->CalculatePlaqutteEnergy(m_fBetaOverN);
}
UBOOL CActionGaugePlaquette::CalculateForceOnGaugeSingleField(const CFieldGaugeSU3* gauge, CFieldGaugeSU3* force) const {
    check(clg_gauge_force(m_pOwner->m_ctx, force->Handle(), gauge->Handle(), 3.0 * m_fBetaOverN), "gauge force");
    return true;
}

void CActionFermionWilsonNf2::Initialize(CLatticeData* owner, const CParameters& p) {
    m_pOwner = owner;
    INT id = 2;
    p.FetchValueINT("FieldId", id);
    m_pFerimionField = dynamic_cast<CFieldFermionWilsonSquareSU3*>(owner->GetFieldById(id));
    if (!m_pFerimionField) throw std::runtime_error("CActionFermionWilsonNf2: FieldId is not a Wilson fermion");
}
clg_action CActionFermionWilsonNf2::Describe() const {
    clg_action a;
    std::memset(&a, 0, sizeof(a));
    a.kind = CLG_ACTION_FERMION_WILSON_NF2;
    a.kappa_or_mass = m_pFerimionField->GetKai();
    a.phi = m_pFerimionField->Handle();
    CSLASolver* s = appGetFermionSolver(m_pFerimionField->m_byFieldId);
    if (!s) throw std::runtime_error("CActionFermionWilsonNf2: no solver for the field");
    a.solver = s->m_params;
    a.solver.kind = CLG_WILSON;
    return a;
}
DOUBLE CActionFermionWilsonNf2::EnergySingleField(const CFieldGaugeSU3* gauge) const {
    clg_action a = Describe();
    double e = 0;
    clg_solve_report rep{};
    check(clg_wilson_pf_energy(m_pOwner->m_ctx, &a.solver, a.phi, gauge->Handle(), a.kappa_or_mass, &e, &rep), "energy");
    return e;
}
UBOOL CActionFermionWilsonNf2::CalculateForceOnGaugeSingleField(const CFieldGaugeSU3* gauge,
                                                                CFieldGaugeSU3* force) const {
    clg_action a = Describe();
    clg_solve_report rep{};
    const int st = clg_wilson_force(m_pOwner->m_ctx, &a.solver, force->Handle(), a.phi, gauge->Handle(),
                                    a.kappa_or_mass, nullptr, nullptr, &rep);
    check(st, "Wilson force");
    return st == CLG_OK;
}

void CActionFermionKS::Initialize(CLatticeData* owner, const CParameters& p) {
    m_pOwner = owner;
    INT id = 2;
    p.FetchValueINT("FieldId", id);
    m_pFerimionField = dynamic_cast<CFieldFermionKSSU3*>(owner->GetFieldById(id));
    if (!m_pFerimionField) throw std::runtime_error("CActionFermionKS: FieldId is not a staggered fermion");
}
clg_action CActionFermionKS::Describe() const {
    clg_action a;
    std::memset(&a, 0, sizeof(a));
    a.kind = CLG_ACTION_FERMION_KS;
    a.kappa_or_mass = m_pFerimionField->GetMass();
    a.phi = m_pFerimionField->Handle();
    a.mc = m_pFerimionField->m_rMC.data();
    a.mc_degree = (int)(m_pFerimionField->m_rMC.size() - 1) / 2;
    a.md = m_pFerimionField->m_rMD.data();
    a.md_degree = (int)(m_pFerimionField->m_rMD.size() - 1) / 2;
    CMultiShiftSolver* ms = appGetMultiShiftSolver(m_pFerimionField->m_byFieldId);
    if (!ms) throw std::runtime_error("CActionFermionKS: no multi-shift solver for the field");
    a.solver = ms->m_params;
    a.solver.kind = CLG_KS;
    a.ms_kind = ms->m_kind;
    a.ms_prec = ms->m_prec;
    return a;
}
DOUBLE CActionFermionKS::EnergySingleField(const CFieldGaugeSU3* gauge) const {
    clg_action a = Describe();
    double e = 0;
    clg_solve_report rep{};
    check(clg_ks_pf_energy(m_pOwner->m_ctx, &a.solver, a.ms_kind, CLG_F64, a.phi, gauge->Handle(), a.kappa_or_mass, a.md,
                           a.md_degree, &e, &rep),
          "KS energy");
    return e;
}
UBOOL CActionFermionKS::CalculateForceOnGaugeSingleField(const CFieldGaugeSU3* gauge, CFieldGaugeSU3* force) const {
    clg_action a = Describe();
    clg_solve_report rep{};
    const int st = clg_ks_force(m_pOwner->m_ctx, &a.solver, a.ms_kind, a.ms_prec, force->Handle(), a.phi,
                                gauge->Handle(), a.kappa_or_mass, a.md, a.md_degree, &rep);
    check(st, "KS force");
    return st == CLG_OK;
}

// CAction::Energy / CalculateForce (CAction.h:42-58): the action's gauge field is the SU(3) field of id 1
namespace {
INT gauge_index(INT gn, const CFieldGauge* const* gf) {
    for (INT i = 0; i < gn; ++i)
        if (gf[i] && gf[i]->m_byFieldId == 1 && dynamic_cast<const CFieldGaugeSU3*>(gf[i])) return i;
    throw std::runtime_error("no SU(3) gauge field of id 1 among the gauge fields");
}
}  // namespace
DOUBLE CAction::Energy(UBOOL, INT gn, INT, const CFieldGauge* const* gf, const CFieldBoson* const*,
                       const CFieldGauge* const*) {
    return EnergySingleField(static_cast<const CFieldGaugeSU3*>(gf[gauge_index(gn, gf)]));
}
UBOOL CAction::CalculateForce(INT gn, INT, const CFieldGauge* const* gf, const CFieldBoson* const*,
                              CFieldGauge* const* gforce, CFieldBoson* const*, CFieldGauge* const*, ESolverPhase) const {
    const INT i = gauge_index(gn, gf);
    CFieldGaugeSU3* F = dynamic_cast<CFieldGaugeSU3*>(gforce[i]);
    if (!F) throw std::runtime_error("CalculateForce: the force field must be an SU(3) gauge field");
    return CalculateForceOnGaugeSingleField(static_cast<const CFieldGaugeSU3*>(gf[i]), F);
}

// ================================================================ measurement, updator
void CMeasurePlaqutteEnergy::OnConfigurationAccepted(const CFieldGaugeSU3* gauge) {
    const CLatticeData* L = gauge->Lattice();
    const double V = (double)L->m_dims[0] * L->m_dims[1] * L->m_dims[2] * L->m_dims[3];
    // sum (1 - ReTr U_p / 3) / (6 V), then 1 - that (CMeasurePlaqutteEnergy.cpp:18-60)
    const double e = gauge->CalculatePlaqutteEnergy(1.0 / 3.0) / (6.0 * V);
    m_lstData.push_back(1.0 - e);
}
Real CMeasurePlaqutteEnergy::GetAverageRealRes() const {
    if (m_lstData.empty()) return 0.0;
    double s = 0;
    for (double v : m_lstData) s += v;
    return s / (double)m_lstData.size();
}

// Updator keys (CHMC / CIntegrator::Initial, CIntegrator.cpp:50-128, CLGLibManager.cpp:282-289)
void CHMC::Initialize(CLatticeData* owner, const CParameters& p) {
    m_pOwner = owner;
    clg_hmc_params_default(&m_params);
    std::string s;
    if (p.FetchStringValue("IntegratorType", s)) {
        if (s == "CIntegratorLeapFrog") m_params.integrator = CLG_INT_LEAPFROG;
        else if (s == "CIntegratorNestedForceGradient") m_params.integrator = CLG_INT_NESTED_FORCE_GRADIENT;
        else throw std::runtime_error("unsupported IntegratorType " + s);
    }
    INT iv;
    Real fv;
    if (p.FetchValueReal("IntegratorStepLength", fv)) m_params.step_length = fv;
    if (p.FetchValueINT("IntegratorStep", iv)) m_params.steps = iv;
    if (p.FetchValueINT("NestedStep", iv)) m_params.nested_steps = iv;
    if (p.FetchValueINT("InnerLeapfrog", iv)) m_params.inner_leapfrog = iv;
    if (p.FetchValueINT("Metropolis", iv)) m_params.metropolis = iv;
    m_uiSeed = owner->m_uiSeed;
}
UINT CHMC::Update(UINT iSteps, UBOOL bMeasure) {
    std::vector<clg_action> acts;
    for (auto& a : m_pOwner->m_pActionList) acts.push_back(a->Describe());
    for (UINT i = 0; i < iSteps; ++i) {
        clg_hmc_report rep{};
        // the reference measures H_diff in its test mode but then accepts every trajectory, and
        // records H_diff only in that mode (CHMC.cpp:100-140)
        clg_hmc_params hp = m_params;
        if (m_bTestHDiff) hp.metropolis = 0;
        check(clg_hmc_trajectory(m_pOwner->m_ctx, m_pOwner->GetGaugeSU3()->Handle(), &hp, acts.data(), (int)acts.size(),
                                 m_uiSeed * 1000003ull + (++m_uiTrajectory), &rep),
              "HMC trajectory");
        m_fLastHDiff = rep.h_diff;
        m_lstSeconds.push_back(rep.seconds);
        if (m_bTestHDiff) {
            m_lstHDiff.push_back(rep.h_diff);
            m_lstH.push_back(rep.h_before);
        }
        const bool accepted = rep.accepted != 0;
        if (accepted) ++m_iAcceptedConfigurationCount;
        std::fprintf(stderr, " HMC: step = %u, H_diff = %f (%f - %f) %s\n", i + 1, rep.h_diff, rep.h_before, rep.h_after,
                     accepted ? "accepted" : "rejected");
        if (bMeasure && accepted && m_pOwner->m_pMeasure) m_pOwner->m_pMeasure->OnConfigurationAccepted(m_pOwner->GetGaugeSU3());
    }
    return m_iAcceptedConfigurationCount;
}
Real CHMC::GetHValue() const {
    if (m_lstH.empty()) return 0.0;
    double s = 0;
    for (double h : m_lstH) s += h;
    return s / (double)m_lstH.size();
}
Real CHMC::GetHDiff() const {
    if (m_lstHDiff.empty()) return 0.0;
    if (m_lstHDiff.size() == 1) return std::fabs(m_lstHDiff[0]);
    double s = 0;
    for (double d : m_lstHDiff) s += d * d;
    return std::sqrt(s / (double)(m_lstHDiff.size() - 1));
}

// ================================================================ lattice and manager
CLatticeData::~CLatticeData() {
    m_pUpdator.reset();
    m_pActionList.clear();
    m_pMSSolvers.clear();
    m_pFermionSolvers.clear();
    m_pPool.clear();
    m_pFermionFields.clear();
    m_pGaugeField.RemoveAll();
    m_pOwnedGauge.reset();
    if (m_ctx) clg_finalize(m_ctx);
}
CField* CLatticeData::GetFieldById(int id) const {
    if (id == 1) return m_pOwnedGauge.get();
    auto it = m_pFermionFields.find(id);
    return it == m_pFermionFields.end() ? nullptr : it->second.get();
}
CField* CLatticeData::GetPooledFieldById(int id) {
    const CField* f = GetFieldById(id);
    if (!f) return nullptr;
    m_pPool.emplace_back(f->GetCopy());
    return m_pPool.back().get();
}
CAction* CLatticeData::GetActionById(int id) const {
    for (const auto& a : m_pActionList)
        if (a->m_byActionId == id) return a.get();
    return nullptr;
}

namespace {
void read_bc(const CParameters& p, int bc[4]) {
    std::vector<INT> period;
    bc[0] = bc[1] = bc[2] = 1;
    bc[3] = -1;
    if (p.FetchValueArrayINT("Period", period) && period.size() == 4)
        for (int i = 0; i < 4; ++i) bc[i] = period[i];
}
template <class F>
void for_each_numbered(const CParameters& sec, const std::string& stem, int count, F&& f) {
    for (int i = 1; i <= std::max(count, 1); ++i) {
        CParameters sub;
        const std::string name = stem + std::to_string(i);
        if (sec.FetchParameterValue(name, sub)) f(sub, i);
        else if (i == 1 && sec.FetchParameterValue(stem, sub)) f(sub, i);
    }
}
}  // namespace

CLatticeData* appInitialCLG(const CParameters& sec, int device) {
    appQuitCLG();
    std::unique_ptr<CLatticeData> L(new CLatticeData);
    try {
        std::vector<INT> len;
        if (!sec.FetchValueArrayINT("LatticeLength", len) || len.size() != 4)
            throw std::runtime_error("LatticeLength [x, y, z, t] is required");
        for (int i = 0; i < 4; ++i) L->m_dims[i] = len[i];
        INT seed = 1234567;
        if (sec.FetchValueINT("RandomSeed", seed)) L->m_uiSeed = (unsigned)seed;
        check(clg_init(device, &L->m_ctx), "clg_init");
        g_lattice = L.get();
        // gauge field (id 1)
        L->m_pOwnedGauge.reset(new CFieldGaugeSU3(L.get()));
        L->m_pOwnedGauge->m_byFieldId = 1;
        L->m_pGaugeField.AddItem(L->m_pOwnedGauge.get());
        {
            CParameters g;
            std::string init = "EFIT_Identity";
            if (sec.FetchParameterValue("Gauge", g)) g.FetchStringValue("FieldInitialType", init);
            if (init == "EFIT_ReadFromFile") {
                // GaugeFileType / GaugeFileName (CLatticeData / CLGLibManager gauge section)
                std::string ftype = "EFFT_CLGBin", fname;
                g.FetchStringValue("GaugeFileType", ftype);
                EFieldFileType t;
                if (!g.FetchStringValue("GaugeFileName", fname) || !appFileTypeFromString(ftype, t))
                    throw std::runtime_error("Gauge: EFIT_ReadFromFile needs GaugeFileName and a known GaugeFileType");
                L->m_pOwnedGauge->InitialFieldWithFile(fname, t);
            } else {
                L->m_pOwnedGauge->InitialField(init == "EFIT_Random" ? EFIT_Random : EFIT_Identity, L->m_uiSeed, 3.0);
            }
        }
        // fermion fields
        INT nf = 1;
        sec.FetchValueINT("FermionFieldCount", nf);
        // even-parity fields (CFieldFermionWilsonSU3DEven) named by a full field's EvenFieldId: a
        // solver registered for one solves its full field's system even-odd preconditioned
        // (CFieldFermionWilsonSquareSU3.cu InverseD with m_byEvenFieldId)
        std::map<int, int> even_parent;
        for_each_numbered(sec, "FermionField", nf, [&](const CParameters& f, int) {
            std::string name;
            f.FetchStringValue("FieldName", name);
            INT id = 2;
            f.FetchValueINT("FieldId", id);
            INT even_id = -1;
            if (f.FetchValueINT("EvenFieldId", even_id)) even_parent[even_id] = id;
            if (name == "CFieldFermionWilsonSU3DEven") return;
            int bc[4];
            read_bc(f, bc);
            std::unique_ptr<CFieldFermion> field;
            if (name == "CFieldFermionWilsonSquareSU3") {
                Real kappa = 0.1;
                f.FetchValueReal("Hopping", kappa);
                field.reset(new CFieldFermionWilsonSquareSU3(L.get(), kappa, bc));
            } else if (name == "CFieldFermionKSSU3") {
                Real mass = 0.1;
                f.FetchValueReal("Mass", mass);
                CFieldFermionKSSU3* ks = new CFieldFermionKSSU3(L.get(), mass, bc);
                f.FetchValueArrayReal("MC", ks->m_rMC);
                f.FetchValueArrayReal("MD", ks->m_rMD);
                field.reset(ks);
            } else {
                throw std::runtime_error("unsupported fermion FieldName " + name);
            }
            field->m_byFieldId = id;
            std::string init;
            if (f.FetchStringValue("FieldInitialType", init) && init == "EFIT_RandomGaussian")
                field->InitialField(EFIT_RandomGaussian, L->m_uiSeed + (unsigned)id);
            L->m_pFermionFields[id] = std::move(field);
        });
        // solvers: Solver / SolverN and MSSolver / MSSolverN, each with SolverForFieldId
        for (const std::string& cn : sec.ChildNames()) {
            CParameters s;
            sec.FetchParameterValue(cn, s);
            std::string name;
            if (!s.FetchStringValue("SolverName", name)) continue;
            INT id = 2;
            s.FetchValueINT("SolverForFieldId", id);
            if (cn.rfind("MSSolver", 0) == 0) {
                std::unique_ptr<CMultiShiftSolver> ms;
                if (name == "CMultiShiftFOM") ms.reset(new CMultiShiftFOM);
                else if (name == "CMultiShiftBiCGStab") ms.reset(new CMultiShiftBiCGStab);
                else if (name == "CMultiShiftGMRES") ms.reset(new CMultiShiftGMRES);
                else if (name == "CMultiShiftNested") ms.reset(new CMultiShiftNested);
                else throw std::runtime_error("unsupported multi-shift solver " + name);
                ms->Configurate(s);
                L->m_pMSSolvers[id] = std::move(ms);
            } else if (cn.rfind("Solver", 0) == 0) {
                std::unique_ptr<CSLASolver> sv;
                if (name == "CSLASolverGMRES") sv.reset(new CSLASolverGMRES);
                else if (name == "CSLASolverBiCGStab") sv.reset(new CSLASolverBiCGStab);
                else if (name == "CSLASolverGCRODR") sv.reset(new CSLASolverGCRODR);
                else throw std::runtime_error("unsupported solver " + name);
                sv->Configurate(s);
                const auto ep = even_parent.find(id);
                if (ep != even_parent.end()) {
                    if (name == "CSLASolverGCRODR")
                        throw std::runtime_error("even-odd GCRO-DR through an even field is not supported");
                    sv->m_params.even_odd = 1;
                    id = ep->second;
                }
                L->m_pFermionSolvers[id] = std::move(sv);
            }
        }
        // actions
        INT na = 0;
        sec.FetchValueINT("ActionListLength", na);
        for_each_numbered(sec, "Action", na, [&](const CParameters& a, int i) {
            std::string name;
            a.FetchStringValue("ActionName", name);
            std::unique_ptr<CAction> act;
            if (name == "CActionGaugePlaquette") act.reset(new CActionGaugePlaquette);
            else if (name == "CActionFermionWilsonNf2") act.reset(new CActionFermionWilsonNf2);
            else if (name == "CActionFermionKS") act.reset(new CActionFermionKS);
            else throw std::runtime_error("unsupported action " + name);
            act->m_byActionId = i;
            act->Initialize(L.get(), a);
            L->m_pActionList.push_back(std::move(act));
        });
        // updator
        CParameters u;
        if (sec.FetchParameterValue("Updator", u)) {
            std::string t;
            u.FetchStringValue("UpdatorType", t);
            if (t != "CHMC") throw std::runtime_error("unsupported UpdatorType " + t);
            L->m_pUpdator.reset(new CHMC);
            L->m_pUpdator->Initialize(L.get(), u);
            INT ep = 0;
            if (sec.FetchValueINT("ExponentialPrecision", ep)) L->m_pUpdator->m_params.exp_precision = ep;
        }
        // measurement
        INT nm = 0;
        sec.FetchValueINT("MeasureListLength", nm);
        for_each_numbered(sec, "Measure", nm, [&](const CParameters& m, int) {
            std::string name;
            m.FetchStringValue("MeasureName", name);
            if (name == "CMeasurePlaqutteEnergy") L->m_pMeasure.reset(new CMeasurePlaqutteEnergy);
        });
    } catch (const std::exception& e) {
        std::fprintf(stderr, "[CLGLib] appInitialCLG: %s\n", e.what());
        g_lattice = nullptr;
        return nullptr;
    }
    return L.release();
}
CLatticeData* appGetLattice() { return g_lattice; }
CSLASolver* appGetFermionSolver(int id) {
    if (!g_lattice) return nullptr;
    auto it = g_lattice->m_pFermionSolvers.find(id);
    return it == g_lattice->m_pFermionSolvers.end() ? nullptr : it->second.get();
}
CMultiShiftSolver* appGetMultiShiftSolver(int id) {
    if (!g_lattice) return nullptr;
    auto it = g_lattice->m_pMSSolvers.find(id);
    return it == g_lattice->m_pMSSolvers.end() ? nullptr : it->second.get();
}
void appQuitCLG() {
    delete g_lattice;
    g_lattice = nullptr;
}

}  // namespace CLGLib
