// clg_test_runner.cpp -- runs CLGTest sections of the reference's YAML test suite through the
// CLGLib facade, in the reference tests' own calling pattern (fields and solvers called with
// _FIELDS, objects reached through appGetLattice()), and prints one JSON line with the quantities
// the reference test checks.
//
//   clg_test_runner <yaml file> <section> <test>
//   test = update   TestUpdateCommon (TestUpdator.cpp:12-140): BeforeMetropolis thermalisation
//                   trajectories, Metropolis measured ones; average plaquette, acceptance, H_diff
//          update_l TestFermionUpdatorL (TestFermionUpdator.cpp:160-219): one trajectory, then Updates-1
//                   measured in the H_diff test mode; plaquette, H_diff, average H, seconds per trajectory
//          solver   TestSolver (TestSolver.cpp:12-121): |D^-1 D phi - phi|^2 and the other three
//                   combinations for D and D^dagger, and DD^dagger
//          mssolver_ks  TestMultiShiftSolverKS (TestMultiShiftSolver.cpp:109-200): rational powers
//                   D^{+-1/2}, D^{+-1/4} and of DD^dagger through the section's multi-shift solver
//          fileio   TestFileIO (TestConfigurationFileIO.cpp:12-37): average plaquette of the gauge
//                   field the section loads (EFIT_ReadFromFile), against ExpectedRes
//          fileio_double / fileio_float  TestFileDOUBLE / TestFileFloat (TestConfigurationFileIO.cpp:
//                   101-162): Dot(U, U) of the reference's 8^4 configuration read in double or float
//          fileio_save  TestFileIOCLG (TestConfigurationFileIO.cpp:55-99): gauge field and fermion field 2
//                   saved and read back, <U', U> / (V Dir) and |psi' - psi|^2
//          fileio_compressed  TestFileIOCLGCompressed (TestConfigurationFileIO.cpp:133-160): the gauge
//                   field saved compressed and read back, |U' - U|^2
//          operators    TestOperators (TestCommon.cpp:13-88): Gaussian-source norm estimators and
//                   <f2, D f1> = <D^dagger f2, f1>
//          gamma5   TestGamma5Hermiticity (TestCommon.cpp:155-168): the dense D checked for gamma5 D
//                   Hermitian (GAMM5Test 0) or D^dagger = (D)^dagger (GAMM5Test 1)
//          antihermitian  TestAnitiHermiticity (TestCommon.cpp:170-180): the dense massless staggered D
//                   anti-Hermitian
//          gauge_invariance  TestGaugeInvarience (TestCommon.cpp:206-239): every action's energy before
//                   and after a random gauge transformation of the gauge and fermion fields
//
// Sections that name a special-frame field or action or a non-torus boundary (rotating or accelerated
// frames, Dirichlet or projective-plane boundaries: none is on the fermion-matrix path) run with the
// standard class of the same family and the torus boundary; the JSON line lists every substitution.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "CLGLib.h"

using namespace CLGLib;

static std::vector<std::string> g_subst;

// replace special-frame class names by the standard ones (recursively) and drop boundary keys
static void substitute(CParameters& p, const std::string& where) {
    static const std::pair<const char*, const char*> names[] = {
        {"CFieldFermionWilsonSquareSU3DR", "CFieldFermionWilsonSquareSU3"},
        {"CFieldFermionWilsonSquareSU3Gamma", "CFieldFermionWilsonSquareSU3"},
        {"CFieldFermionKSSU3R", "CFieldFermionKSSU3"},
        {"CFieldFermionKSSU3RigidAcc", "CFieldFermionKSSU3"},
        {"CFieldFermionKSSU3EM", "CFieldFermionKSSU3"},
        {"CFieldFermionKSSU3DR", "CFieldFermionKSSU3"},
        {"CFieldFermionKSSU3Gamma", "CFieldFermionKSSU3"},
        {"CFieldGaugeSU3D", "CFieldGaugeSU3"},
        {"CActionGaugePlaquetteRotating", "CActionGaugePlaquette"},
        {"CActionGaugePlaquetteAcceleration", "CActionGaugePlaquette"},
    };
    for (const char* key : {"FieldName", "ActionName"}) {
        std::string v;
        if (!p.FetchStringValue(key, v)) continue;
        for (const auto& n : names)
            if (v == n.first) {
                p.SetStringValue(key, n.second);
                g_subst.push_back(where + "." + key + ": " + n.first + " -> " + n.second);
            }
    }
    std::string b;
    if (p.FetchStringValue("LatticeBoundary", b) && b != "CBoundaryConditionTorusSquare") {
        p.SetStringValue("LatticeBoundary", "CBoundaryConditionTorusSquare");
        g_subst.push_back(where + ".LatticeBoundary: " + b + " -> CBoundaryConditionTorusSquare");
    }
    for (const std::string& c : p.ChildNames()) substitute(p.Child(c), where.empty() ? c : where + "." + c);
}

static std::string subst_json() {
    std::string s = "[";
    for (size_t i = 0; i < g_subst.size(); ++i) s += (i ? ", \"" : "\"") + g_subst[i] + "\"";
    return s + "]";
}

static int run_update(CParameters& sec) {
    CLatticeData* L = appGetLattice();
    std::vector<Real> expect;
    sec.FetchValueArrayReal("ExpectedRes", expect);
    INT before = 3, metro = 12, miss = 2;
    sec.FetchValueINT("BeforeMetropolis", before);
    sec.FetchValueINT("Metropolis", metro);
    sec.FetchValueINT("ExpectMiss", miss);
    if (!L->m_pUpdator || !L->m_pMeasure) {
        std::fprintf(stderr, "section has no updator or measurement\n");
        return 2;
    }
    const auto t0 = std::chrono::steady_clock::now();
    L->m_pUpdator->Update((UINT)before, false);
    L->m_pMeasure->Reset();
    L->m_pUpdator->SetTestHdiff(true);
    L->m_pUpdator->Update((UINT)metro, true);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const UINT accepted = L->m_pUpdator->GetConfigurationCount() - (UINT)before;
    std::printf("{\"test\": \"update\", \"plaquette_avg\": %.8f, \"plaquette_last\": %.8f, \"expected\": %.8f, "
                "\"accepted\": %u, \"metropolis\": %d, \"expect_accept\": %d, \"hdiff\": %.6f, \"last_hdiff\": %.6f, "
                "\"seconds\": %.3f, \"substituted\": %s}\n",
                L->m_pMeasure->GetAverageRealRes(), L->m_pMeasure->GetLastRealRes(), expect.empty() ? 0.0 : expect[0],
                accepted, metro, metro - miss, L->m_pUpdator->GetHDiff(), std::fabs(L->m_pUpdator->m_fLastHDiff), secs,
                subst_json().c_str());
    return 0;
}

static int run_update_l(CParameters& sec) {
    CLatticeData* L = appGetLattice();
    Real expected = 0.17;
    sec.FetchValueReal("ExpectedRes", expected);
    INT updates = 3;
    sec.FetchValueINT("Updates", updates);
    if (updates > 10 || updates < 3) updates = 3;
    if (!L->m_pUpdator || !L->m_pMeasure) return 2;
    L->m_pUpdator->Update(1, false);
    L->m_pUpdator->SetTestHdiff(true);
    L->m_pMeasure->Reset();
    L->m_pUpdator->Update((UINT)(updates - 1), true);
    const std::vector<double>& t = L->m_pUpdator->m_lstSeconds;
    std::string ts = "[";
    for (size_t i = 0; i < t.size(); ++i) ts += (i ? ", " : "") + std::to_string(t[i]);
    ts += "]";
    std::printf("{\"test\": \"update_l\", \"plaquette_avg\": %.8f, \"expected\": %.8f, \"hdiff\": %.8f, \"h\": %.6f, "
                "\"trajectories\": %d, \"seconds_per_trajectory\": %s, \"substituted\": %s}\n",
                L->m_pMeasure->GetAverageRealRes(), expected, L->m_pUpdator->GetHDiff(), L->m_pUpdator->GetHValue(),
                updates, ts.c_str(), subst_json().c_str());
    return 0;
}

// TestSolver's pattern: each |op^-1 op phi - phi|^2 and |op op^-1 phi - phi|^2 on a copy of field 2
static int run_solver(CParameters& sec) {
    Real expected = 1e-4;
    sec.FetchValueReal("ExpectedErr", expected);
    CFieldFermion* phi = dynamic_cast<CFieldFermion*>(appGetLattice()->GetFieldById(2));
    if (!phi) return 2;
    double err[6];
    int k = 0;
    auto finish = [&](CFieldFermion* r) {
        r->AxpyMinus(phi);
        err[k++] = _cuCabsf(r->DotReal(r));
        delete r;
    };
    const EFieldOperator fwd[3] = {EFO_F_D, EFO_F_Ddagger, EFO_F_DDdagger};
    const EFieldOperator inv[3] = {EFO_F_InverseD, EFO_F_InverseDdagger, EFO_F_InverseDDdagger};
    for (int o = 0; o < 3; ++o) {
        CFieldFermion* r = dynamic_cast<CFieldFermion*>(phi->GetCopy());
        r->ApplyOperator(fwd[o], _FIELDS);
        r->ApplyOperator(inv[o], _FIELDS);
        finish(r);
        r = dynamic_cast<CFieldFermion*>(phi->GetCopy());
        r->ApplyOperator(inv[o], _FIELDS);
        r->ApplyOperator(fwd[o], _FIELDS);
        finish(r);
    }
    std::printf("{\"test\": \"solver\", \"errors\": [%.6e, %.6e, %.6e, %.6e, %.6e, %.6e], \"expected\": %.3e, "
                "\"phi2\": %.6f, \"substituted\": %s}\n",
                err[0], err[1], err[2], err[3], err[4], err[5], expected, phi->DotReal(phi).x, subst_json().c_str());
    return 0;
}

static int run_mssolver_ks(CParameters& sec) {
    // TestMultiShiftSolverKS (TestMultiShiftSolver.cpp:109-200): |R^-1/n R^1/n phi - phi|^2 for n = 2, 4
    // with op = D and DD^dagger, the rational coefficients taken from the section
    std::vector<Real> c2, m2, c4, m4;
    sec.FetchValueArrayReal("OneOver2", c2);
    sec.FetchValueArrayReal("MinusOneOver2", m2);
    sec.FetchValueArrayReal("OneOver4", c4);
    sec.FetchValueArrayReal("MinusOneOver4", m4);
    Real expected = 1e-5;
    sec.FetchValueReal("ExpectedErr", expected);
    CFieldFermion* phi = dynamic_cast<CFieldFermion*>(appGetLattice()->GetFieldById(2));
    if (!phi) return 2;
    const CRatinalApproximation r2(c2), r_2(m2), r4(c4), r_4(m4);
    double err[4];
    int k = 0;
    for (EFieldOperator op : {EFO_F_D, EFO_F_DDdagger})
        for (const auto* pr : {&r2, &r4}) {
            std::unique_ptr<CFieldFermion> t(dynamic_cast<CFieldFermion*>(phi->GetCopy()));
            t->RationalApproximation(op, _FIELDS, pr);
            t->RationalApproximation(op, _FIELDS, pr == &r2 ? &r_2 : &r_4);
            t->AxpyMinus(phi);
            err[k++] = _cuCabsf(t->DotReal(t.get()));
        }
    std::printf("{\"test\": \"mssolver_ks\", \"errors\": [%.6e, %.6e, %.6e, %.6e], \"expected\": %.3e, "
                "\"phi2\": %.6f}\n",
                err[0], err[1], err[2], err[3], expected, phi->DotReal(phi).x);
    return 0;
}

static int run_fileio(CParameters& sec) {
    CLatticeData* L = appGetLattice();
    Real expected = 0.625129946974942;
    sec.FetchValueReal("ExpectedRes", expected);
    const double V = (double)L->m_dims[0] * L->m_dims[1] * L->m_dims[2] * L->m_dims[3];
    const CFieldGaugeSU3* g = dynamic_cast<const CFieldGaugeSU3*>(L->m_pGaugeField[0]);
    const double plaq = 1.0 - g->CalculatePlaqutteEnergy(1.0 / 3.0) / (6.0 * V);
    std::printf("{\"test\": \"fileio\", \"plaquette\": %.15f, \"expected\": %.15f}\n", plaq, expected);
    return 0;
}

// TestFileDOUBLE / TestFileFloat (TestConfigurationFileIO.cpp:101-162): the reference's 8^4 gauge
// configuration in double or float read into the lattice's gauge field; Dot(U, U).x (49152 expected)
static int run_fileio_dot(bool dbl) {
    CLatticeData* L = appGetLattice();
    CField* g = L->m_pGaugeField[0];
    if (!g) return 2;
    const std::string file = std::string("../Debug/") + (dbl ? "testGaugeDouble.con_" : "testGaugeSingle.con_");
    g->InitialFieldWithFile(file, dbl ? EFFT_CLGBinDouble : EFFT_CLGBinFloat);
    const CLGComplex r = g->Dot(g);
    std::printf("{\"test\": \"%s\", \"dot\": [%.10f, %.10f], \"expected\": 49152.0}\n",
                dbl ? "fileio_double" : "fileio_float", r.x, r.y);
    return 0;
}

// TestFileIOCLG's pattern (TestConfigurationFileIO.cpp:55-99): save the gauge field and fermion field 2
// in the CLG binary format, read them into new fields; <U', U> / (V Dir) and |psi' - psi|^2
static int run_fileio_save(CParameters&) {
    CLatticeData* L = appGetLattice();
    const CField* g = L->m_pGaugeField[0];
    const CField* f = L->GetFieldById(2);
    if (!g || !f) return 2;
    const std::string md5g = g->SaveToFile("testGauge.con");
    const std::string md5f = f->SaveToFile("testFermion.con");
    CField* ng = g->GetCopy();
    CField* nf = f->GetCopy();
    ng->Zero();
    nf->Zero();
    ng->InitialFieldWithFile("testGauge.con", EFFT_CLGBin);
    nf->InitialFieldWithFile("testFermion.con", EFFT_CLGBin);
    const double V = (double)L->m_dims[0] * L->m_dims[1] * L->m_dims[2] * L->m_dims[3];
    const CLGComplex r1 = ng->DotReal(g);
    nf->AxpyMinus(f);
    const CLGComplex r2 = nf->DotReal(nf);
    delete ng;
    delete nf;
    std::printf("{\"test\": \"fileio_save\", \"gauge\": [%.12f, %.12f], \"fermion\": [%.6e, %.6e], "
                "\"md5_gauge\": \"%s\", \"md5_fermion\": \"%s\"}\n",
                r1.x / (V * 4), r1.y / (V * 4), r2.x, r2.y, md5g.c_str(), md5f.c_str());
    return 0;
}

// TestFileIOCLGCompressed's pattern (TestConfigurationFileIO.cpp:133-160): save the gauge field
// compressed, read it into a new field, |U' - U|^2
static int run_fileio_compressed(CParameters&) {
    CLatticeData* L = appGetLattice();
    const CFieldGaugeSU3* g = dynamic_cast<const CFieldGaugeSU3*>(L->m_pGaugeField[0]);
    if (!g) return 2;
    const auto t0 = std::chrono::steady_clock::now();
    const std::string md5 = g->SaveToCompressedFile("testGaugeCompressed.con");
    CFieldGaugeSU3* n = new CFieldGaugeSU3(L);
    n->InitialFieldWithFile("testGaugeCompressed.con", EFFT_CLGBinCompressed);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    n->AxpyMinus(g);
    const CLGComplex res = n->DotReal(n);
    delete n;
    std::printf("{\"test\": \"fileio_compressed\", \"res\": [%.6e, %.6e], \"md5\": \"%s\", \"seconds\": %.3f}\n",
                res.x, res.y, md5.c_str(), secs);
    return 0;
}

static int run_operators(CParameters&) {
    CLatticeData* L = appGetLattice();
    const double V = (double)L->m_dims[0] * L->m_dims[1] * L->m_dims[2] * L->m_dims[3];
    CFieldFermionWilsonSquareSU3* f1 = dynamic_cast<CFieldFermionWilsonSquareSU3*>(L->GetPooledFieldById(2));
    CFieldFermionWilsonSquareSU3* f2 = dynamic_cast<CFieldFermionWilsonSquareSU3*>(L->GetPooledFieldById(2));
    CFieldFermionWilsonSquareSU3* f3 = dynamic_cast<CFieldFermionWilsonSquareSU3*>(L->GetPooledFieldById(2));
    CFieldFermionWilsonSquareSU3* f4 = dynamic_cast<CFieldFermionWilsonSquareSU3*>(L->GetPooledFieldById(2));
    if (!f1 || !f2 || !f3 || !f4) return 2;
    f1->InitialField(EFIT_RandomGaussian);
    f2->InitialField(EFIT_RandomGaussian);
    const CLGComplex d11 = f1->DotReal(f1), d22 = f2->DotReal(f2), d12 = f1->DotReal(f2), d21 = f2->DotReal(f1);
    f1->CopyTo(f3);
    f3->D(_FIELDS);
    f2->CopyTo(f4);
    f4->Ddagger(_FIELDS);
    const CLGComplex d5 = f2->DotReal(f3), d6 = f4->DotReal(f1);
    const double n = 12 * V;
    std::printf("{\"test\": \"operators\", \"f1f1\": [%.8f, %.8f], \"f2f2\": [%.8f, %.8f], \"f1f2\": [%.8f, %.8f], "
                "\"f2f1\": [%.8f, %.8f], \"f2_Df1\": [%.12f, %.12f], \"Ddag_f2_f1\": [%.12f, %.12f], \"n\": %.0f, "
                "\"substituted\": %s}\n",
                d11.x / n, d11.y, d22.x / n, d22.y, d12.x / n, d12.y / n, d21.x / n, d21.y / n, d5.x, d5.y, d6.x, d6.y,
                n, subst_json().c_str());
    f1->Return();
    f2->Return();
    f3->Return();
    f4->Return();
    return 0;
}

static int run_gamma5(CParameters& sec) {
    INT iG5 = 0;
    sec.FetchValueINT("GAMM5Test", iG5);
    CLatticeData* L = appGetLattice();
    CFieldFermionWilsonSquareSU3* f = dynamic_cast<CFieldFermionWilsonSquareSU3*>(L->GetPooledFieldById(2));
    if (!f) return 2;
    const UINT errors = f->TestGamma5Hermitian(L->m_pGaugeField[0], iG5 != 0);
    const double e0 = f->m_fLastTestError;
    // the other form as well, so both identities are reported
    const UINT errors2 = f->TestGamma5Hermitian(L->m_pGaugeField[0], iG5 == 0);
    std::printf("{\"test\": \"gamma5\", \"gamma5_test\": %d, \"errors\": %u, \"max_rel_dev\": %.3e, "
                "\"other_errors\": %u, \"other_max_rel_dev\": %.3e, \"substituted\": %s}\n",
                iG5, errors, e0, errors2, f->m_fLastTestError, subst_json().c_str());
    f->Return();
    return 0;
}

static int run_antihermitian(CParameters&) {
    CLatticeData* L = appGetLattice();
    CFieldFermionKSSU3* f = dynamic_cast<CFieldFermionKSSU3*>(L->GetPooledFieldById(2));
    if (!f) return 2;
    const UINT errors = f->TestAntiHermitian(_FIELDS);
    std::printf("{\"test\": \"antihermitian\", \"errors\": %u, \"max_rel_dev\": %.3e, \"substituted\": %s}\n", errors,
                f->m_fLastTestError, subst_json().c_str());
    f->Return();
    return 0;
}

static int run_gauge_invariance(CParameters&) {
    CLatticeData* L = appGetLattice();
    const INT na = (INT)L->m_pActionList.size();
    std::vector<double> before, after;
    for (INT i = 0; i < na; ++i) before.push_back(L->GetActionById(i + 1)->Energy(false, _FIELDS, nullptr));
    CGaugeFixingRandom random(L);
    // the transformation really changes the gauge field
    std::unique_ptr<CField> copy(L->m_pGaugeField[0]->GetCopy());
    random.GaugeFixing(L->m_pGaugeField[0]);
    copy->AxpyMinus(L->m_pGaugeField[0]);
    const double moved = copy->DotReal(copy.get()).x;
    for (auto& kv : L->m_pFermionFields) random.AlsoFixingFermion(kv.second.get());
    for (INT i = 0; i < na; ++i) after.push_back(L->GetActionById(i + 1)->Energy(false, _FIELDS, nullptr));
    std::string b = "[", a = "[";
    for (INT i = 0; i < na; ++i) {
        char buf[64];
        std::snprintf(buf, sizeof(buf), "%s%.15e", i ? ", " : "", before[(size_t)i]);
        b += buf;
        std::snprintf(buf, sizeof(buf), "%s%.15e", i ? ", " : "", after[(size_t)i]);
        a += buf;
    }
    std::printf("{\"test\": \"gauge_invariance\", \"before\": %s], \"after\": %s], \"gauge_moved\": %.6e, "
                "\"substituted\": %s}\n",
                b.c_str(), a.c_str(), moved, subst_json().c_str());
    return 0;
}

static void dump(const CParameters& p) {
    std::printf("{");
    bool first = true;
    for (auto& kv : p.Values()) {
        std::string v;
        for (char c : kv.second) {
            if (c == '"' || c == '\\') v += '\\';
            v += c;
        }
        std::printf("%s\"%s\": \"%s\"", first ? "" : ", ", kv.first.c_str(), v.c_str());
        first = false;
    }
    for (auto& kv : p.Children()) {
        std::printf("%s\"%s\": ", first ? "" : ", ", kv.first.c_str());
        dump(kv.second);
        first = false;
    }
    std::printf("}");
}

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr,
                     "usage: %s <yaml> <section> <update|update_l|solver|mssolver_ks|fileio|fileio_double|fileio_float|fileio_save|fileio_compressed|"
                     "operators|gamma5|"
                     "antihermitian|gauge_invariance|dump>\n",
                     argv[0]);
        return 2;
    }
    CParameters all, sec;
    if (!CYAMLParser::ParseFile(argv[1], all) || !all.FetchParameterValue(argv[2], sec)) {
        std::fprintf(stderr, "cannot read section %s of %s\n", argv[2], argv[1]);
        return 2;
    }
    const std::string t = argv[3];
    if (t == "dump") {  // the parsed section, no device needed
        dump(sec);
        std::printf("\n");
        return 0;
    }
    substitute(sec, "");
    if (!appInitialCLG(sec)) return 3;
    int rc = 2;
    try {
        if (t == "update") rc = run_update(sec);
        else if (t == "update_l") rc = run_update_l(sec);
        else if (t == "solver") rc = run_solver(sec);
        else if (t == "fileio") rc = run_fileio(sec);
        else if (t == "fileio_compressed") rc = run_fileio_compressed(sec);
        else if (t == "fileio_save") rc = run_fileio_save(sec);
        else if (t == "fileio_double" || t == "fileio_float") rc = run_fileio_dot(t == "fileio_double");
        else if (t == "mssolver_ks") rc = run_mssolver_ks(sec);
        else if (t == "operators") rc = run_operators(sec);
        else if (t == "gamma5") rc = run_gamma5(sec);
        else if (t == "antihermitian") rc = run_antihermitian(sec);
        else if (t == "gauge_invariance") rc = run_gauge_invariance(sec);
        else std::fprintf(stderr, "unknown test %s\n", argv[3]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "failed: %s\n", e.what());
        rc = 4;
    }
    appQuitCLG();
    return rc;
}
