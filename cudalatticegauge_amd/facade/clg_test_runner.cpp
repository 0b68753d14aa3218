// clg_test_runner.cpp -- runs CLGTest sections of the reference's YAML test suite through the
// CLGLib facade and prints one JSON line with the quantities the reference test checks.
//
//   clg_test_runner <yaml file> <section> <test>
//   test = update   TestUpdateCommon (TestUpdator.cpp:12-140): BeforeMetropolis thermalisation
//                   trajectories, Metropolis measured ones; average plaquette, acceptance, H_diff
//          solver   TestSolver (TestSolver.cpp:12-121): |D^-1 D phi - phi|^2 and the other three
//                   combinations for D and D^dagger, and DD^dagger
//          mssolver_ks  TestMultiShiftSolverKS (TestMultiShiftSolver.cpp:109-200): rational powers
//                   D^{+-1/2}, D^{+-1/4} and of DD^dagger through the section's multi-shift solver
//          fileio   TestFileIO (TestConfigurationFileIO.cpp:12-37): average plaquette of the gauge
//                   field the section loads (EFIT_ReadFromFile), against ExpectedRes
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>

#include "CLGLib.h"

using namespace CLGLib;

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
                "\"seconds\": %.3f}\n",
                L->m_pMeasure->GetAverageRealRes(), L->m_pMeasure->GetLastRealRes(), expect.empty() ? 0.0 : expect[0],
                accepted, metro, metro - miss, L->m_pUpdator->GetHDiff(), std::fabs(L->m_pUpdator->m_fLastHDiff), secs);
    return 0;
}

static int run_solver(CParameters& sec) {
    CLatticeData* L = appGetLattice();
    Real expected = 1e-4;
    sec.FetchValueReal("ExpectedErr", expected);
    CFieldFermion* phi = dynamic_cast<CFieldFermion*>(L->GetFieldById(2));
    const CFieldGaugeSU3* g = L->m_pGaugeField.get();
    if (!phi) return 2;
    double err[5];
    auto diff = [&](CField* r) {
        r->AxpyMinus(phi);
        return std::abs(r->Dot(r));
    };
    std::unique_ptr<CFieldFermion> r(dynamic_cast<CFieldFermion*>(phi->GetCopy()));
    r->D(g);
    r->ApplyOperator(EFO_F_InverseD, g);
    err[0] = diff(r.get());
    r.reset(dynamic_cast<CFieldFermion*>(phi->GetCopy()));
    r->ApplyOperator(EFO_F_InverseD, g);
    r->D(g);
    err[1] = diff(r.get());
    r.reset(dynamic_cast<CFieldFermion*>(phi->GetCopy()));
    r->Ddagger(g);
    r->ApplyOperator(EFO_F_InverseDdagger, g);
    err[2] = diff(r.get());
    r.reset(dynamic_cast<CFieldFermion*>(phi->GetCopy()));
    r->ApplyOperator(EFO_F_InverseDdagger, g);
    r->Ddagger(g);
    err[3] = diff(r.get());
    r.reset(dynamic_cast<CFieldFermion*>(phi->GetCopy()));
    r->DDdagger(g);
    r->ApplyOperator(EFO_F_InverseDDdagger, g);
    err[4] = diff(r.get());
    std::printf("{\"test\": \"solver\", \"errors\": [%.6e, %.6e, %.6e, %.6e, %.6e], \"expected\": %.3e, "
                "\"phi2\": %.6f}\n",
                err[0], err[1], err[2], err[3], err[4], expected, std::abs(phi->Dot(phi)));
    return 0;
}

static int run_mssolver_ks(CParameters& sec) {
    // TestMultiShiftSolverKS (TestMultiShiftSolver.cpp:109-200): |R^-1/n R^1/n phi - phi|^2 for n = 2, 4
    // with op = D and DD^dagger, the rational coefficients taken from the section
    CLatticeData* L = appGetLattice();
    std::vector<Real> c2, m2, c4, m4;
    sec.FetchValueArrayReal("OneOver2", c2);
    sec.FetchValueArrayReal("MinusOneOver2", m2);
    sec.FetchValueArrayReal("OneOver4", c4);
    sec.FetchValueArrayReal("MinusOneOver4", m4);
    Real expected = 1e-5;
    sec.FetchValueReal("ExpectedErr", expected);
    CFieldFermion* phi = dynamic_cast<CFieldFermion*>(L->GetFieldById(2));
    const CFieldGaugeSU3* g = L->m_pGaugeField.get();
    if (!phi) return 2;
    const CRatinalApproximation r2(c2), r_2(m2), r4(c4), r_4(m4);
    double err[4];
    int k = 0;
    for (EFieldOperator op : {EFO_F_D, EFO_F_DDdagger})
        for (const auto* pr : {&r2, &r4}) {
            std::unique_ptr<CFieldFermion> t(dynamic_cast<CFieldFermion*>(phi->GetCopy()));
            t->RationalApproximation(op, g, pr);
            t->RationalApproximation(op, g, pr == &r2 ? &r_2 : &r_4);
            t->AxpyMinus(phi);
            err[k++] = std::abs(t->Dot(t.get()));
        }
    std::printf("{\"test\": \"mssolver_ks\", \"errors\": [%.6e, %.6e, %.6e, %.6e], \"expected\": %.3e, "
                "\"phi2\": %.6f}\n",
                err[0], err[1], err[2], err[3], expected, std::abs(phi->Dot(phi)));
    return 0;
}

static int run_fileio(CParameters& sec) {
    CLatticeData* L = appGetLattice();
    Real expected = 0.625129946974942;
    sec.FetchValueReal("ExpectedRes", expected);
    const double V = (double)L->m_dims[0] * L->m_dims[1] * L->m_dims[2] * L->m_dims[3];
    const double plaq = 1.0 - L->m_pGaugeField->CalculatePlaqutteEnergy(1.0 / 3.0) / (6.0 * V);
    std::printf("{\"test\": \"fileio\", \"plaquette\": %.15f, \"expected\": %.15f}\n", plaq, expected);
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
        std::fprintf(stderr, "usage: %s <yaml> <section> <update|solver|mssolver_ks|fileio|dump>\n", argv[0]);
        return 2;
    }
    CParameters all, sec;
    if (!CYAMLParser::ParseFile(argv[1], all) || !all.FetchParameterValue(argv[2], sec)) {
        std::fprintf(stderr, "cannot read section %s of %s\n", argv[2], argv[1]);
        return 2;
    }
    if (std::string(argv[3]) == "dump") {  // the parsed section, no device needed
        dump(sec);
        std::printf("\n");
        return 0;
    }
    if (!appInitialCLG(sec)) return 3;
    int rc = 2;
    try {
        const std::string t = argv[3];
        if (t == "update") rc = run_update(sec);
        else if (t == "solver") rc = run_solver(sec);
        else if (t == "fileio") rc = run_fileio(sec);
        else if (t == "mssolver_ks") rc = run_mssolver_ks(sec);
        else std::fprintf(stderr, "unknown test %s\n", argv[3]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "failed: %s\n", e.what());
        rc = 4;
    }
    appQuitCLG();
    return rc;
}
