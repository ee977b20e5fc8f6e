// CBCTForwardProj.h -- polyenergetic projector base class with the interface
// of the reference's CBCTForwardProj (Source/CBCTForwardProj.h:10-100).
//
// The reference loops over energy bins on the host, re-uploading and
// re-marching on the GPU per bin.  Here one call to the C ABI (ctfp_forward)
// marches every material once per ray and applies the whole spectrum; the
// per-energy host steps keep their names for code that calls them directly.
// Failures throw ctfp::Error instead of exiting the process.
#pragma once
#include <string>
#include <vector>

#include "DataType.h"
#include "Grid.h"
#include "PhantomObject.h"

class CBCTForwardProj {
  public:
    CBCTForwardProj();
    virtual ~CBCTForwardProj();

    virtual void computePolyForwProj() = 0;
    virtual void testI0() = 0;

    virtual void computeParas();
    virtual void initGridInfo();

    virtual void addIandIAbsorb();
    virtual void convertIandIAbsorbDt();
    virtual void saveIandIAbsorb();

    virtual void addStandardI0(float &I0Standard);
    virtual void addI0();
    virtual void saveI0();

    virtual void computePerScinThinckness();
    virtual void updateDetResponse();
    virtual void computeFoSpOffset();

    virtual void readSpecrtumNorm();
    virtual void readGridMassAttu();
    virtual void readScintillatorMassAttu();
    virtual void getADetResponse();
    virtual void saveADetResponse();

    virtual void showScanSystemInfo();
    virtual void showGridmInfo();
    virtual void showProcessInfo();

    virtual void computeProj();
    virtual void saveProj();

    virtual void creatOutputFolder();
    virtual void _removeOutputFolder();

  public:
    PhantomObject *mPhantomObject = nullptr;
    CTScanSystemInfo mCTScanSystemInfo;
    CTScanParas mCTScanParas;
    PolyForwardProj h_mForwardProj;
    FilePath mFilePath;
    PhantomMaterial _mPhantomMaterial;
    GridInfo mGridInfo;
    GridAndDetectorSystem mGridAndDetectorSystem;
    std::vector<size_t> gridPeriodIndex;

    size_t specIndex = 0;
    float *I = nullptr;
    uint32 *IPolyenergetic = nullptr;
    float *IAbsorb = nullptr;
    uint32 *IPolyenergeticAbsorb = nullptr;
    float *aDetResponse = nullptr;

    // Options with no reference counterpart.
    int device = 0;                      // HIP device (the reference uses GPU 0)
    bool cleanOutputFolder = false;      // delete the output folder first, as the reference does
    bool honorInterspaceDensity = false; // the reference's grid drivers drop InterspaceDensity
    bool verbose = true;
    float lastI0Standard = 0; // I0 at the nominal scintillator thickness (testI0)

  protected:
    // Runs one scan through the C ABI and writes the reference's output files.
    void runScan(bool singleMaterial, bool withGrid, bool withResponse, bool focalSpot = false);
    void runTestI0(bool withGrid);
    void allocHost(bool withGrid);
    std::vector<float> gridTable() ; // T[e][u] for every bin
    virtual const char *tag() const = 0; // "Grid" / "NoGrid"
    bool hasGrid_ = false;
    bool hasResponse_ = true;

    // storage behind the public raw pointers
    std::vector<float> vI, vIA, vI0, vProj, vSpec, vGridT, vGridMu, vInterMu, vScinMu, vThick,
        vDet, vGridRow, vADet, vFoSpU, vFoSpV;
    std::vector<uint32> vIu, vIAu;
};
