// CBCTPolyForwardProjGrid.h -- multi-material polyenergetic projector, with the
// focused grid; interface of the reference class (Source/CBCTPolyForwardProjGrid.h:11-44).
#pragma once
#include "CBCTForwardProj.h"

class CBCTPolyForwardProjGrid : public CBCTForwardProj {
  public:
    CBCTPolyForwardProjGrid(CTScanParas inCTScanParas, GridInfo inGridInfo, FilePath inFileInfo, const PhantomMaterial inPhantomMaterial);
    ~CBCTPolyForwardProjGrid();

    // Label volume -> per-material densities, all energy bins, with response.
    virtual void computePolyForwProj();
    // Same without detector response (the reference's run() left this
    // combination empty, ForwardProjection.cpp:60-64,72-75).
    virtual void computePolyForwProjNoResponse();
    // Flat field only, host side (no GPU).
    virtual void testI0();

  protected:
    const char *tag() const override { return "Grid"; }
};
