// CBCTSinMatPolyForwardProjGrid.h -- single-material projector with grid;
// interface of the reference class (Source/CBCTSinMatPolyForwardProjGrid.h:5-21).
// The density is label * Density1 (PhantomObject::computeDesity).
#pragma once
#include "CBCTPolyForwardProjGrid.h"

class CBCTSinMatPolyForwardProjGrid : private CBCTPolyForwardProjGrid {
  public:
    CBCTSinMatPolyForwardProjGrid(CTScanParas inCTScanParas, GridInfo inGridInfo, FilePath inFileInfo,
                                  const PhantomMaterial inphantomMaterial)
        : CBCTPolyForwardProjGrid(inCTScanParas, inGridInfo, inFileInfo, inphantomMaterial) {}
    ~CBCTSinMatPolyForwardProjGrid();

    virtual void computePolyForwProj();
    void computePolyForwProjNoResponse();
    // The reference's focal-spot variant (unreachable from its run() and
    // broken there): a point source replaced by per-view source offsets.
    void computePolyForwProjFoSp();

    using CBCTForwardProj::I;
    using CBCTForwardProj::IAbsorb;
    using CBCTForwardProj::IPolyenergetic;
    using CBCTForwardProj::IPolyenergeticAbsorb;
    using CBCTForwardProj::aDetResponse;
    using CBCTForwardProj::h_mForwardProj;
    using CBCTForwardProj::mCTScanSystemInfo;
    using CBCTForwardProj::device;
    using CBCTForwardProj::cleanOutputFolder;
    using CBCTForwardProj::honorInterspaceDensity;
    using CBCTForwardProj::verbose;
    using CBCTForwardProj::gridPeriodIndex;
};
