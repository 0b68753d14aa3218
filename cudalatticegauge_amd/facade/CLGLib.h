// CLGLib.h -- C++ facade with CLGLib's class names over the C ABI of libclg_hip.so.
//
// Code written against CLGLib's field / solver / action / updator classes (CField.h:89-247,
// CFieldFermion.h:16-325, CSLASolver.h:24-60, CMultiShiftSolver.h:24-50, CAction.h:18-109,
// CUpdator.h:25-110) and CLGTest's YAML sections keeps its names and call pattern: the objects
// are created by class name from a parameter tree (appCreate + CLGLibManager, CLGLibManager.cpp:
// 513-578, 749-775) and reached through appGetLattice / appGetFermionSolver / appGetMultiShiftSolver.
// Only what the fermion-matrix path and its HMC need is reproduced; every numerical call goes to
// the HIP library.
#pragma once

#include <complex>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/clg_hip.h"

namespace CLGLib {

typedef int INT;
typedef unsigned int UINT;
typedef double Real;
typedef double DOUBLE;
typedef bool UBOOL;
typedef std::complex<double> CLGComplex;

// EFieldOperator (CField.h:37-56) and EOperatorCoefficientType (CField.h:62-68): the C ABI's
// enumerations, same names and numbering
typedef clg_field_operator EFieldOperator;
typedef clg_coeff_type EOperatorCoefficientType;
// ESolverPhase (CCommonData.h:168-172)
enum ESolverPhase { ESP_StartTrajectory = 0, ESP_InTrajectory = 1, ESP_EndTrajectory = 2, ESP_Once = 3 };
// EFieldInitialType subset (CFieldGaugeSU3.cu:32-82, CFieldFermionWilsonSquareSU3.cu:463-500,
// CLatticeData.h:18-33)
enum EFieldInitialType { EFIT_Zero, EFIT_Identity, EFIT_Random, EFIT_RandomGaussian, EFIT_ReadFromFile };
// EFieldFileType, same names and numbering (CField.h:26-35)
enum EFieldFileType {
    EFFT_BridgePPTXT = CLG_FILE_BRIDGEPP_TXT,
    EFFT_BridgePPBin = CLG_FILE_BRIDGEPP_BIN,
    EFFT_CLGBin = CLG_FILE_CLG_BIN,
    EFFT_CLGBinCompressed = CLG_FILE_CLG_BIN_COMPRESSED,
    EFFT_CLGBinFloat = CLG_FILE_CLG_BIN_FLOAT,
    EFFT_CLGBinDouble = CLG_FILE_CLG_BIN_DOUBLE
};
// "EFFT_BridgePPTXT" -> EFFT_BridgePPTXT, ... (the YAML key GaugeFileType); false when unknown
bool appFileTypeFromString(const std::string& name, EFieldFileType& out);

// ---------------------------------------------------------------- parameters (YAML tree)
class CParameters {
public:
    UBOOL FetchValueINT(const std::string& key, INT& v) const;
    UBOOL FetchValueReal(const std::string& key, Real& v) const;
    UBOOL FetchStringValue(const std::string& key, std::string& v) const;
    UBOOL FetchValueArrayReal(const std::string& key, std::vector<Real>& v) const;
    UBOOL FetchValueArrayINT(const std::string& key, std::vector<INT>& v) const;
    UBOOL FetchParameterValue(const std::string& key, CParameters& sub) const;
    UBOOL Exist(const std::string& key) const;
    void SetStringValue(const std::string& key, const std::string& v) { m_values[key] = v; }
    CParameters& Child(const std::string& key) { return m_children[key]; }
    std::vector<std::string> ChildNames() const;
    const std::map<std::string, std::string>& Values() const { return m_values; }
    const std::map<std::string, CParameters>& Children() const { return m_children; }

private:
    std::map<std::string, std::string> m_values;
    std::map<std::string, CParameters> m_children;
};

// The YAML subset of the reference's configuration files: `key : value` lines, nested maps by
// indentation, inline arrays `[a, b]`, `#` comments (CYAMLParser, Tools/CYAMLParser.h).
class CYAMLParser {
public:
    static UBOOL ParseFile(const std::string& path, CParameters& out);
    static UBOOL ParseText(const std::string& text, CParameters& out);
};

class CLatticeData;

// ---------------------------------------------------------------- fields
// CRatinalApproximation (CRationalApproximation.h:38-117): coefficients [c, a_1..a_n, b_1..b_n]
class CRatinalApproximation {
public:
    explicit CRatinalApproximation(const std::vector<Real>& coeffs) : m_c(coeffs) {}
    int Degree() const { return (int)(m_c.size() - 1) / 2; }
    std::vector<Real> m_c;
};

class CField {
public:
    virtual ~CField();
    clg_field Handle() const { return m_h; }
    CLatticeData* Lattice() const { return m_pOwner; }
    void Zero();
    void CopyTo(CField* dst) const;
    void AxpyPlus(const CField* x);
    void AxpyMinus(const CField* x);
    void Axpy(Real a, const CField* x);
    void Axpy(const CLGComplex& a, const CField* x);
    void ScalarMultply(const CLGComplex& a);
    void ScalarMultply(Real a) { ScalarMultply(CLGComplex(a, 0.0)); }
    CLGComplex Dot(const CField* other) const;
    CLGComplex DotReal(const CField* other) const { return Dot(other); }
    virtual CField* GetCopy() const = 0;
    std::vector<double> Download() const;
    void Upload(const std::vector<double>& host);
    // CField::InitialFieldWithFile / SaveToFile (CField.h:97, 154; CField.cpp:32-72); SaveToFile
    // returns the MD5 string of the bytes written
    void InitialFieldWithFile(const std::string& fileName, EFieldFileType type = EFFT_CLGBin);
    std::string SaveToFile(const std::string& fileName, EFieldFileType type = EFFT_CLGBin) const;

    int m_byFieldId = 0;

protected:
    CField(CLatticeData* owner, int kind);
    CLatticeData* m_pOwner;
    clg_field m_h = nullptr;
    int m_kind;
};

class CFieldGaugeSU3 : public CField {
public:
    explicit CFieldGaugeSU3(CLatticeData* owner);
    void InitialField(EFieldInitialType t, unsigned seed = 1, Real eps = 3.0);
    // Sum over plaquettes of (1 - ReTr U_p / 3) times 3 * betaOverN (CFieldGaugeSU3.cu:308-371)
    DOUBLE CalculatePlaqutteEnergy(DOUBLE betaOverN) const;
    CField* GetCopy() const override;
};

class CFieldFermion : public CField {
public:
    // D, Ddagger, DD, DDdagger applied in place with an EOperatorCoefficientType factor
    // (CFieldFermion.h:46-241)
    virtual void D(const CFieldGaugeSU3* gauge, int eCoeffType = EOCT_None, Real re = 1.0, Real im = 0.0);
    virtual void Ddagger(const CFieldGaugeSU3* gauge, int eCoeffType = EOCT_None, Real re = 1.0, Real im = 0.0);
    virtual void DD(const CFieldGaugeSU3* gauge, int eCoeffType = EOCT_None, Real re = 1.0, Real im = 0.0);
    virtual void DDdagger(const CFieldGaugeSU3* gauge, int eCoeffType = EOCT_None, Real re = 1.0, Real im = 0.0);
    void ApplyOperator(EFieldOperator op, const CFieldGaugeSU3* gauge, int eCoeffType = EOCT_None, Real re = 1.0,
                       Real im = 0.0);
    // through the solver registered for this field id (appGetFermionSolver)
    UBOOL InverseD(const CFieldGaugeSU3* gauge);
    UBOOL InverseDdagger(const CFieldGaugeSU3* gauge);
    UBOOL InverseDD(const CFieldGaugeSU3* gauge);
    UBOOL InverseDDdagger(const CFieldGaugeSU3* gauge);
    void InitialField(EFieldInitialType t, unsigned seed = 2);
    // EFS_Point source at (site coordinates, spin, colour) (CFieldFermionWilsonSquareSU3.cu:1110-1160)
    void InitialAsSource(const int coord[4], int spin, int colour);
    virtual void PrepareForHMC(const CFieldGaugeSU3* gauge, unsigned seed) = 0;
    // CFieldFermion::RationalApproximation (CField.cpp:114-148): this <- c this + sum_k a_k (op + b_k)^-1 this,
    // through the multi-shift solver registered for this field id
    void RationalApproximation(EFieldOperator op, const CFieldGaugeSU3* gauge, const class CRatinalApproximation* r);
    Real Param() const { return m_fParam; }

protected:
    CFieldFermion(CLatticeData* owner, int kind, Real param, const int bc[4]);
    virtual int Kind() const = 0;
    UBOOL Inverse(EFieldOperator op, const CFieldGaugeSU3* gauge);
    Real m_fParam;
    int m_bc[4];
};

class CFieldFermionWilsonSquareSU3 : public CFieldFermion {
public:
    CFieldFermionWilsonSquareSU3(CLatticeData* owner, Real kappa, const int bc[4]);
    void PrepareForHMC(const CFieldGaugeSU3* gauge, unsigned seed) override;
    CField* GetCopy() const override;
    Real GetKai() const { return m_fParam; }

protected:
    int Kind() const override { return CLG_WILSON; }
};

class CFieldFermionKSSU3 : public CFieldFermion {
public:
    CFieldFermionKSSU3(CLatticeData* owner, Real mass, const int bc[4]);
    void PrepareForHMC(const CFieldGaugeSU3* gauge, unsigned seed) override;
    CField* GetCopy() const override;
    Real GetMass() const { return m_fParam; }
    // rational approximations [c, a_1..a_n, b_1..b_n] (YAML MC / MD, CField.cpp:164-194)
    std::vector<Real> m_rMC, m_rMD;
    void D_MC(const CFieldGaugeSU3* gauge);
    void D_MD(const CFieldGaugeSU3* gauge);

protected:
    int Kind() const override { return CLG_KS; }
};

// ---------------------------------------------------------------- solvers
class CSLASolver {
public:
    virtual ~CSLASolver() {}
    virtual void Configurate(const CParameters& param);
    // x = op^-1 b (op one of the EFO_F_Inverse* operators or the operator to invert)
    virtual UBOOL Solve(CField* pFieldX, const CField* pFieldB, const CFieldGaugeSU3* gauge, EFieldOperator uiM,
                        ESolverPhase ePhase = ESP_Once, const CField* pStart = nullptr);
    clg_solver_params m_params{};
    clg_solve_report m_lastReport{};
};
class CSLASolverGMRES : public CSLASolver {
public:
    CSLASolverGMRES();
};
class CSLASolverBiCGStab : public CSLASolver {
public:
    CSLASolverBiCGStab();
};
class CSLASolverGCRODR : public CSLASolver {
public:
    CSLASolverGCRODR();
    ~CSLASolverGCRODR() override;
    void Configurate(const CParameters& param) override;
    UBOOL Solve(CField* pFieldX, const CField* pFieldB, const CFieldGaugeSU3* gauge, EFieldOperator uiM,
                ESolverPhase ePhase = ESP_Once, const CField* pStart = nullptr) override;
    clg_gcrodr_params m_gparams{};

private:
    clg_gcrodr m_solver = nullptr;
    clg_context m_ctx = nullptr;
};

class CMultiShiftSolver {
public:
    explicit CMultiShiftSolver(int kind);
    virtual ~CMultiShiftSolver() {}
    virtual void Configurate(const CParameters& param);
    // x_n = (op + c_n)^-1 b for every shift (CMultiShiftSolver.h:38-47)
    virtual UBOOL Solve(std::vector<CField*>& pFieldX, const std::vector<CLGComplex>& cn, const CField* pFieldB,
                        const CFieldGaugeSU3* gauge, EFieldOperator uiM);
    clg_solver_params m_params{};
    int m_kind;
    int m_prec = CLG_F64;
};
class CMultiShiftFOM : public CMultiShiftSolver {
public:
    CMultiShiftFOM() : CMultiShiftSolver(CLG_MS_FOM) {}
};
class CMultiShiftBiCGStab : public CMultiShiftSolver {
public:
    CMultiShiftBiCGStab() : CMultiShiftSolver(CLG_MS_BICGSTAB) {}
};
class CMultiShiftGMRES : public CMultiShiftSolver {
public:
    CMultiShiftGMRES() : CMultiShiftSolver(CLG_MS_GMRES) {}
};
class CMultiShiftNested : public CMultiShiftSolver {
public:
    CMultiShiftNested() : CMultiShiftSolver(CLG_MS_NESTED) {}
};

// ---------------------------------------------------------------- actions, updator, measurement
class CAction {
public:
    virtual ~CAction() {}
    virtual UBOOL IsFermion() const { return false; }
    virtual void Initialize(CLatticeData* owner, const CParameters& param) = 0;
    virtual clg_action Describe() const = 0;
    virtual DOUBLE Energy(const CFieldGaugeSU3* gauge) = 0;
    virtual void CalculateForce(const CFieldGaugeSU3* gauge, CFieldGaugeSU3* force) = 0;
    int m_byActionId = 0;

protected:
    CLatticeData* m_pOwner = nullptr;
};
class CActionGaugePlaquette : public CAction {
public:
    void Initialize(CLatticeData* owner, const CParameters& param) override;
    clg_action Describe() const override;
    DOUBLE Energy(const CFieldGaugeSU3* gauge) override;
    void CalculateForce(const CFieldGaugeSU3* gauge, CFieldGaugeSU3* force) override;
    Real m_fBetaOverN = 5.0 / 3.0;
};
class CActionFermionWilsonNf2 : public CAction {
public:
    UBOOL IsFermion() const override { return true; }
    void Initialize(CLatticeData* owner, const CParameters& param) override;
    clg_action Describe() const override;
    DOUBLE Energy(const CFieldGaugeSU3* gauge) override;
    void CalculateForce(const CFieldGaugeSU3* gauge, CFieldGaugeSU3* force) override;
    CFieldFermionWilsonSquareSU3* m_pFerimionField = nullptr;
};
class CActionFermionKS : public CAction {
public:
    UBOOL IsFermion() const override { return true; }
    void Initialize(CLatticeData* owner, const CParameters& param) override;
    clg_action Describe() const override;
    DOUBLE Energy(const CFieldGaugeSU3* gauge) override;
    void CalculateForce(const CFieldGaugeSU3* gauge, CFieldGaugeSU3* force) override;
    CFieldFermionKSSU3* m_pFerimionField = nullptr;
};

class CMeasurePlaqutteEnergy {
public:
    // <ReTr U_p> / 3 of an accepted configuration (CMeasurePlaqutteEnergy.cpp:18-60)
    void OnConfigurationAccepted(const CFieldGaugeSU3* gauge);
    void Reset() { m_lstData.clear(); }
    Real GetLastRealRes() const { return m_lstData.empty() ? 0.0 : m_lstData.back(); }
    Real GetAverageRealRes() const;
    std::vector<Real> m_lstData;
};

class CHMC {
public:
    void Initialize(CLatticeData* owner, const CParameters& param);
    // CHMC::Update (CHMC.cpp:76-219): iSteps trajectories; the gauge field is replaced on acceptance
    UINT Update(UINT iSteps, UBOOL bMeasure);
    void SetTestHdiff(UBOOL b) { m_bTestHDiff = b; }
    void SetAutoCorrection(UBOOL) {}
    UINT GetConfigurationCount() const { return m_iAcceptedConfigurationCount; }
    // sqrt(sum dH^2 / (n - 1)) over the recorded trajectories (CUpdator.h:66-86)
    Real GetHDiff() const;
    Real m_fLastHDiff = 0;
    clg_hmc_params m_params{};
    std::vector<double> m_lstHDiff;

private:
    CLatticeData* m_pOwner = nullptr;
    UBOOL m_bTestHDiff = false;
    UINT m_iAcceptedConfigurationCount = 0;
    unsigned long long m_uiTrajectory = 0;
    unsigned long long m_uiSeed = 1;
};

// ---------------------------------------------------------------- the lattice (CLatticeData + manager)
class CLatticeData {
public:
    ~CLatticeData();
    clg_context m_ctx = nullptr;
    int m_dims[4] = {0, 0, 0, 0};
    std::unique_ptr<CFieldGaugeSU3> m_pGaugeField;
    std::map<int, std::unique_ptr<CFieldFermion>> m_pFermionFields;  // by FieldId
    std::map<int, std::unique_ptr<CSLASolver>> m_pFermionSolvers;    // SolverForFieldId
    std::map<int, std::unique_ptr<CMultiShiftSolver>> m_pMSSolvers;  // SolverForFieldId
    std::vector<std::unique_ptr<CAction>> m_pActionList;
    std::unique_ptr<CHMC> m_pUpdator;
    std::unique_ptr<CMeasurePlaqutteEnergy> m_pMeasure;
    CField* GetFieldById(int id) const;
    unsigned m_uiSeed = 1234567;
};

// build a lattice from a CLGTest section (Dim, LatticeLength, Gauge, FermionFieldN, Solver, MSSolver,
// ActionN, Updator); returns nullptr and prints the reason on failure
CLatticeData* appInitialCLG(const CParameters& section, int device = 0);
CLatticeData* appGetLattice();
CSLASolver* appGetFermionSolver(int fieldId);
CMultiShiftSolver* appGetMultiShiftSolver(int fieldId);
void appQuitCLG();

}  // namespace CLGLib
