// CBCTSinMatPolyForwardProjNoGrid.h -- single-material projector without grid;
// interface of the reference class (Source/CBCTSinMatPolyForwardProjNoGrid.h:3-17).
#pragma once
#include "CBCTPolyForwardProjNoGrid.h"

class CBCTSinMatPolyForwardProjNoGrid : private CBCTPolyForwardProjNoGrid {
  public:
    CBCTSinMatPolyForwardProjNoGrid(CTScanParas inCTScanParas, FilePath inFileInfo,
                                    const PhantomMaterial inPhantomMaterial)
        : CBCTPolyForwardProjNoGrid(inCTScanParas, inFileInfo, inPhantomMaterial) {}
    ~CBCTSinMatPolyForwardProjNoGrid();

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
    using CBCTForwardProj::verbose;
};
