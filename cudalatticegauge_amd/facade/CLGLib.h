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

#include <cmath>
#include <initializer_list>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/clg_hip.h"

namespace CLGLib {

typedef int INT;
typedef unsigned int UINT;
typedef unsigned char BYTE;
typedef double Real;
typedef double DOUBLE;
typedef bool UBOOL;
// the literal helpers CLGLib code is written with: F(v) a Real literal (CLGFloat.h:40, the
// double-precision build) and _T(s) a narrow string (PlatformDefine.h:134)
#ifndef F
#define F(v) (v)
#endif
#ifndef _T
#define _T(s) s
#endif

// CLGComplex / cuDoubleComplex: two doubles .x (real) and .y (imaginary), as the reference's type
// (CLGLib_Private.h, cuComplex.h), with its constructor and modulus helpers
struct CLGComplex {
    double x = 0.0, y = 0.0;
};
typedef CLGComplex cuDoubleComplex;
inline CLGComplex _make_cuComplex(double re, double im) { return CLGComplex{re, im}; }
inline double _cuCabsf(const CLGComplex& z) { return std::sqrt(z.x * z.x + z.y * z.y); }
inline double _cuCrealf(const CLGComplex& z) { return z.x; }
inline double _cuCimagf(const CLGComplex& z) { return z.y; }

// TArray (Tools/TArray.h): the reference's growable array; Num() / GetData() feed the
// (count, pointer) argument pairs of the field and solver calls
template <class T>
class TArray {
public:
    TArray() {}
    TArray(std::initializer_list<T> l) : m_v(l) {}
    INT Num() const { return (INT)m_v.size(); }
    T* GetData() { return m_v.data(); }
    const T* GetData() const { return m_v.data(); }
    INT AddItem(const T& v) {
        m_v.push_back(v);
        return (INT)m_v.size() - 1;
    }
    void RemoveAll() { m_v.clear(); }
    T& operator[](INT i) { return m_v[(size_t)i]; }
    const T& operator[](INT i) const { return m_v[(size_t)i]; }
    typename std::vector<T>::iterator begin() { return m_v.begin(); }
    typename std::vector<T>::iterator end() { return m_v.end(); }
    typename std::vector<T>::const_iterator begin() const { return m_v.begin(); }
    typename std::vector<T>::const_iterator end() const { return m_v.end(); }

private:
    std::vector<T> m_v;
};

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
    void ScalarMultply(Real a) { ScalarMultply(_make_cuComplex(a, 0.0)); }
    CLGComplex Dot(const CField* other) const;
    CLGComplex DotReal(const CField* other) const { return Dot(other); }
    virtual CField* GetCopy() const = 0;
    // give a pooled field back (CField::Return); the lattice keeps pooled fields until it is destroyed
    void Return() {}
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

// CFieldGauge (CFieldGauge.h): the base of every gauge field; the field and solver calls take
// arrays of these (gaugeNum, gaugeFields) and pick the one whose id the fermion field names
class CFieldGauge : public CField {
protected:
    CFieldGauge(CLatticeData* owner, int kind) : CField(owner, kind) {}
};
// CFieldBoson (CFieldBoson.h): no boson field is on the fermion-matrix path; the calls accept the
// reference's (bosonNum, bosonFields) pair and ignore it
class CFieldBoson {
public:
    virtual ~CFieldBoson() {}
};

class CFieldGaugeSU3 : public CFieldGauge {
public:
    explicit CFieldGaugeSU3(CLatticeData* owner);
    void InitialField(EFieldInitialType t, unsigned seed = 1, Real eps = 3.0);
    // Sum over plaquettes of (1 - ReTr U_p / 3) times 3 * betaOverN (CFieldGaugeSU3.cu:308-371)
    DOUBLE CalculatePlaqutteEnergy(DOUBLE betaOverN) const;
    CField* GetCopy() const override;
    // the per-link logarithms, 9 reals a link (CFieldGaugeSU3.cu:1651-1690); returns the MD5 string;
    // InitialFieldWithFile(name, EFFT_CLGBinCompressed) reads them back
    std::string SaveToCompressedFile(const std::string& fileName) const;
};

// The fermion calls keep the reference's argument lists (CFieldFermion.h:20-246): the gauge and boson
// fields come as (count, array) pairs, which CLGTest passes as _FIELDS.
class CFieldFermion : public CField {
public:
    // this <- coeff * op(this) (CFieldFermion.h:160-200)
    virtual void D(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields, const CFieldBoson* const* pBoson,
                   EOperatorCoefficientType eCoeffType = EOCT_None, Real fCoeffReal = 1.0, Real fCoeffImg = 0.0);
    virtual void Ddagger(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                         const CFieldBoson* const* pBoson, EOperatorCoefficientType eCoeffType = EOCT_None,
                         Real fCoeffReal = 1.0, Real fCoeffImg = 0.0);
    virtual void DD(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields, const CFieldBoson* const* pBoson,
                    EOperatorCoefficientType eCoeffType = EOCT_None, Real fCoeffReal = 1.0, Real fCoeffImg = 0.0);
    virtual void DDdagger(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                          const CFieldBoson* const* pBoson, EOperatorCoefficientType eCoeffType = EOCT_None,
                          Real fCoeffReal = 1.0, Real fCoeffImg = 0.0);
    // the same with the mass (staggered 2am) or hopping parameter (Wilson) replaced by fMass
    // (CFieldFermion.h:201-244)
    virtual void DWithMass(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                           const CFieldBoson* const* pBoson, Real fMass, EOperatorCoefficientType eCoeffType = EOCT_None,
                           Real fCoeffReal = 1.0, Real fCoeffImg = 0.0);
    virtual void DdaggerWithMass(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                                 const CFieldBoson* const* pBoson, Real fMass,
                                 EOperatorCoefficientType eCoeffType = EOCT_None, Real fCoeffReal = 1.0,
                                 Real fCoeffImg = 0.0);
    virtual void DDWithMass(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                            const CFieldBoson* const* pBoson, Real fMass, EOperatorCoefficientType eCoeffType = EOCT_None,
                            Real fCoeffReal = 1.0, Real fCoeffImg = 0.0);
    virtual void DDdaggerWithMass(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                                  const CFieldBoson* const* pBoson, Real fMass,
                                  EOperatorCoefficientType eCoeffType = EOCT_None, Real fCoeffReal = 1.0,
                                  Real fCoeffImg = 0.0);
    // CFieldFermion::ApplyOperator (CFieldFermion.h:43-157): the operators above, the inverses through
    // the field's solver, EFO_F_RationalD with a CRatinalApproximation* and the *_WithMass operators
    // with a Real* mass in pOtherParameters; FALSE for what it does not implement
    UBOOL ApplyOperator(EFieldOperator op, INT gaugeNum, INT bosonNum, const CFieldGauge* const* pGauge,
                        const CFieldBoson* const* pBoson, EOperatorCoefficientType eCoeffType = EOCT_None,
                        Real fCoeffReal = 1.0, Real fCoeffImg = 0.0, void* pOtherParameters = nullptr);
    // through the solver registered for this field id (appGetFermionSolver; CFieldFermion.cpp:267-289)
    virtual UBOOL InverseD(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                           const CFieldBoson* const* pBoson);
    virtual UBOOL InverseDdagger(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                                 const CFieldBoson* const* pBoson);
    virtual UBOOL InverseDD(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                            const CFieldBoson* const* pBoson);
    virtual UBOOL InverseDDdagger(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                                  const CFieldBoson* const* pBoson);
    // CFieldFermion::RationalApproximation (CField.cpp:114-148): this <- c this + sum_k a_k (op + b_k)^-1 this,
    // through the multi-shift solver registered for this field id
    virtual UBOOL RationalApproximation(EFieldOperator op, INT gaugeNum, INT bosonNum,
                                        const CFieldGauge* const* gaugeFields, const CFieldBoson* const* pBoson,
                                        const class CRatinalApproximation* pRational);
    // heat-bath of the pseudofermion (CFieldFermion.h:20-29)
    virtual void PrepareForHMC(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                               const CFieldBoson* const* pBoson);
    // F += the fermion force of this pseudofermion on the gauge field it couples to
    // (CFieldFermion.h:34-43); FALSE when a solve fails
    virtual UBOOL CalculateForce(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                                 const CFieldBoson* const* pBoson, CFieldGauge* const* pGaugeForce,
                                 CFieldBoson* const* pBosonForce, ESolverPhase ePhase) const;
    // EFIT_RandomGaussian draws from the lattice's seed sequence (a new draw per call, as the
    // reference's random number generator does); EFIT_Zero clears the field
    void InitialField(EFieldInitialType t);
    void InitialField(EFieldInitialType t, unsigned seed);
    // EFS_Point source at (site coordinates, spin, colour) (CFieldFermionWilsonSquareSU3.cu:1110-1160)
    void InitialAsSource(const int coord[4], int spin, int colour);
    Real Param() const { return m_fParam; }
    // the id of the gauge field this fermion couples to (m_byGaugeFieldIds[0]; 1 by default)
    BYTE m_byGaugeFieldId = 1;
    // the gauge field of id m_byGaugeFieldId in a (count, array) pair, or nullptr
    const CFieldGaugeSU3* GetDefaultGauge(INT gaugeNum, const CFieldGauge* const* gaugeFields) const;

protected:
    CFieldFermion(CLatticeData* owner, int kind, Real param, const int bc[4]);
    virtual int Kind() const = 0;
    virtual void PrepareForHMCS(const CFieldGaugeSU3* gauge, unsigned seed) = 0;
    void Apply(EFieldOperator op, const CFieldGaugeSU3* gauge, Real param, EOperatorCoefficientType ct, Real re,
               Real im);
    UBOOL Inverse(EFieldOperator op, INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                  const CFieldBoson* const* pBoson);
    Real m_fParam;
    int m_bc[4];
    unsigned m_uiHeatBath = 0;  // heat-bath draws made so far (seeds the next one)
};

class CFieldFermionWilsonSquareSU3 : public CFieldFermion {
public:
    CFieldFermionWilsonSquareSU3(CLatticeData* owner, Real kappa, const int bc[4]);
    CField* GetCopy() const override;
    Real GetKai() const { return m_fParam; }
    // CFieldFermionWilsonSquareSU3::TestGamma5Hermitian (CFieldFermionWilsonSquareSU3.cu:1362-1500): the
    // dense matrix of D from point sources; bTestGamma5 false: count the entries where gamma5 D is not
    // Hermitian, true: where the implemented D^dagger differs from (D)^dagger (tolerance 1e-12 of the
    // largest entry); m_fLastTestError keeps the largest deviation
    UINT TestGamma5Hermitian(const CFieldGauge* pGauge, UBOOL bTestGamma5);
    Real m_fLastTestError = 0;

protected:
    int Kind() const override { return CLG_WILSON; }
    void PrepareForHMCS(const CFieldGaugeSU3* gauge, unsigned seed) override;
};

class CFieldFermionKSSU3 : public CFieldFermion {
public:
    CFieldFermionKSSU3(CLatticeData* owner, Real mass, const int bc[4]);
    CField* GetCopy() const override;
    Real GetMass() const { return m_fParam; }
    // rational approximations [c, a_1..a_n, b_1..b_n] (YAML MC / MD, CField.cpp:164-194)
    std::vector<Real> m_rMC, m_rMD;
    // this <- r_MC(D D^dagger) this, r_MD(D D^dagger) this (CFieldFermionKS.h:78-97)
    void D_MC(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields, const CFieldBoson* const* pBoson);
    void D_MD(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields, const CFieldBoson* const* pBoson);
    // CFieldFermionKS::TestAntiHermitian (CFieldFermionKSSU3.cu:1424-1490): the dense matrix of the massless
    // D from point sources; counts the entries where D_0 + D_0^dagger != 0 (tolerance 1e-12 of the largest
    // entry); m_fLastTestError keeps the largest deviation
    UINT TestAntiHermitian(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                           const CFieldBoson* const* pBoson);
    Real m_fLastTestError = 0;

protected:
    int Kind() const override { return CLG_KS; }
    void PrepareForHMCS(const CFieldGaugeSU3* gauge, unsigned seed) override;
};

// ---------------------------------------------------------------- solvers
class CSLASolver {
public:
    virtual ~CSLASolver() {}
    virtual void Configurate(const CParameters& param);
    // x = uiM^-1 b for uiM the operator (EFO_F_D .. EFO_F_DDdagger) or its EFO_F_Inverse* name, with the
    // reference's argument list (CSLASolver.h:42-50); x may be b
    virtual UBOOL Solve(CField* pFieldX, const CField* pFieldB, INT gaugeNum, INT bosonNum,
                        const CFieldGauge* const* gaugeFields, const CFieldBoson* const* bosonFields,
                        EFieldOperator uiM, ESolverPhase ePhase = ESP_Once, const CField* pStart = nullptr);
    UBOOL IsAbsoluteAccuracy() const { return m_params.absolute_accuracy != 0; }
    clg_solver_params m_params{};
    clg_solve_report m_lastReport{};
    CLatticeData* m_pOwner = nullptr;
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
    UBOOL Solve(CField* pFieldX, const CField* pFieldB, INT gaugeNum, INT bosonNum,
                const CFieldGauge* const* gaugeFields, const CFieldBoson* const* bosonFields, EFieldOperator uiM,
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
    // x_n = (uiM + c_n)^-1 b for every shift, with the reference's argument list
    // (CMultiShiftSolver.h:38-47); x_n must not be b
    virtual UBOOL Solve(TArray<CField*>& pFieldX, const TArray<CLGComplex>& cn, const CField* pFieldB, INT gaugeNum,
                        INT bosonNum, const CFieldGauge* const* gaugeFields, const CFieldBoson* const* bosonFields,
                        EFieldOperator uiM, ESolverPhase ePhase = ESP_Once, const CField* pStart = nullptr);
    UBOOL IsAbsoluteAccuracy() const { return m_params.absolute_accuracy != 0; }
    CLatticeData* m_pOwner = nullptr;
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
// Actions keep the reference's entry points (CAction.h:42-79); the single-gauge-field work is in the
// *SingleField methods, as in the reference's subclasses
class CAction {
public:
    virtual ~CAction() {}
    virtual UBOOL IsFermion() const { return false; }
    virtual void Initialize(CLatticeData* owner, const CParameters& param) = 0;
    virtual clg_action Describe() const = 0;
    // S of the configuration (bBeforeEvolution and the staple fields are accepted and not needed here)
    DOUBLE Energy(UBOOL bBeforeEvolution, INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                  const CFieldBoson* const* bosonFields, const CFieldGauge* const* stableFields);
    // gaugeForces[i] += the force of this action on gaugeFields[i]; FALSE when a solve fails
    UBOOL CalculateForce(INT gaugeNum, INT bosonNum, const CFieldGauge* const* gaugeFields,
                         const CFieldBoson* const* bosonFields, CFieldGauge* const* gaugeForces,
                         CFieldBoson* const* bosonForces, CFieldGauge* const* stapleFields, ESolverPhase ePhase) const;
    BYTE GetActionId() const { return (BYTE)m_byActionId; }
    int m_byActionId = 0;

protected:
    virtual DOUBLE EnergySingleField(const CFieldGaugeSU3* gauge) const = 0;
    virtual UBOOL CalculateForceOnGaugeSingleField(const CFieldGaugeSU3* gauge, CFieldGaugeSU3* force) const = 0;
    CLatticeData* m_pOwner = nullptr;
};
class CActionGaugePlaquette : public CAction {
public:
    void Initialize(CLatticeData* owner, const CParameters& param) override;
    clg_action Describe() const override;
    Real m_fBetaOverN = 5.0 / 3.0;

protected:
    DOUBLE EnergySingleField(const CFieldGaugeSU3* gauge) const override;
    UBOOL CalculateForceOnGaugeSingleField(const CFieldGaugeSU3* gauge, CFieldGaugeSU3* force) const override;
};
class CActionFermionWilsonNf2 : public CAction {
public:
    UBOOL IsFermion() const override { return true; }
    void Initialize(CLatticeData* owner, const CParameters& param) override;
    clg_action Describe() const override;
    CFieldFermionWilsonSquareSU3* m_pFerimionField = nullptr;

protected:
    DOUBLE EnergySingleField(const CFieldGaugeSU3* gauge) const override;
    UBOOL CalculateForceOnGaugeSingleField(const CFieldGaugeSU3* gauge, CFieldGaugeSU3* force) const override;
};
class CActionFermionKS : public CAction {
public:
    UBOOL IsFermion() const override { return true; }
    void Initialize(CLatticeData* owner, const CParameters& param) override;
    clg_action Describe() const override;
    CFieldFermionKSSU3* m_pFerimionField = nullptr;

protected:
    DOUBLE EnergySingleField(const CFieldGaugeSU3* gauge) const override;
    UBOOL CalculateForceOnGaugeSingleField(const CFieldGaugeSU3* gauge, CFieldGaugeSU3* force) const override;
};

// CGaugeFixingRandom (CGaugeFixingRandom.cu): a random local gauge transformation g(x), drawn once,
// applied to the gauge field and then to fermion fields (TestCommon.cpp:206-239)
class CGaugeFixingRandom {
public:
    explicit CGaugeFixingRandom(CLatticeData* owner, Real eps = 3.0);
    ~CGaugeFixingRandom();
    void GaugeFixing(CFieldGauge* pGauge);
    void AlsoFixingFermion(CFieldFermion* pFermion);

private:
    clg_field m_gt = nullptr;
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
    // average H (before the trajectory) over the recorded trajectories (CUpdator.h:89-102)
    Real GetHValue() const;
    Real m_fLastHDiff = 0;
    clg_hmc_params m_params{};
    std::vector<double> m_lstHDiff, m_lstH;
    std::vector<double> m_lstSeconds;  // wall time of each trajectory

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
    // the gauge and boson fields as the reference holds them (CLatticeData.h:14, the _FIELDS pair);
    // the gauge field itself is owned by m_pOwnedGauge
    TArray<CFieldGauge*> m_pGaugeField;
    TArray<CFieldBoson*> m_pBosonField;
    std::unique_ptr<CFieldGaugeSU3> m_pOwnedGauge;
    CFieldGaugeSU3* GetGaugeSU3() const { return m_pOwnedGauge.get(); }
    std::map<int, std::unique_ptr<CFieldFermion>> m_pFermionFields;  // by FieldId
    std::map<int, std::unique_ptr<CSLASolver>> m_pFermionSolvers;    // SolverForFieldId
    std::map<int, std::unique_ptr<CMultiShiftSolver>> m_pMSSolvers;  // SolverForFieldId
    std::vector<std::unique_ptr<CAction>> m_pActionList;
    std::unique_ptr<CHMC> m_pUpdator;
    std::unique_ptr<CMeasurePlaqutteEnergy> m_pMeasure;
    CField* GetFieldById(int id) const;
    // a field of the same class and parameters as field `id`, owned by the lattice (CLatticeData::
    // GetPooledFieldById, CLatticeData.cu:162); its content is not defined until set
    CField* GetPooledFieldById(int id);
    // the action with ActionId id (1-based, CLatticeData::GetActionById)
    CAction* GetActionById(int id) const;
    unsigned m_uiSeed = 1234567;
    // the next seed of the lattice's sequence of random draws
    unsigned NextSeed() { return m_uiSeed * 2654435761u + 97u * (++m_uiDraws); }

private:
    std::vector<std::unique_ptr<CField>> m_pPool;
    unsigned m_uiDraws = 0;
};

// build a lattice from a CLGTest section (Dim, LatticeLength, Gauge, FermionFieldN, Solver, MSSolver,
// ActionN, Updator); returns nullptr and prints the reason on failure
CLatticeData* appInitialCLG(const CParameters& section, int device = 0);
// the (count, array) arguments every field and solver call takes, from the current lattice
// (CLatticeData.h:14)
#define _FIELDS                                                                                           \
    ::CLGLib::appGetLattice()->m_pGaugeField.Num(), ::CLGLib::appGetLattice()->m_pBosonField.Num(),       \
        ::CLGLib::appGetLattice()->m_pGaugeField.GetData(), ::CLGLib::appGetLattice()->m_pBosonField.GetData()
CLatticeData* appGetLattice();
CSLASolver* appGetFermionSolver(int fieldId);
CMultiShiftSolver* appGetMultiShiftSolver(int fieldId);
void appQuitCLG();

}  // namespace CLGLib
