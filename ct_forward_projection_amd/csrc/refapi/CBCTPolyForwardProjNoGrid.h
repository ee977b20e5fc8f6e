// CBCTPolyForwardProjNoGrid.h -- multi-material polyenergetic projector, without the
// focused grid; interface of the reference class (Source/CBCTPolyForwardProjNoGrid.h:9-36).
#pragma once
#include "CBCTForwardProj.h"

class CBCTPolyForwardProjNoGrid : public CBCTForwardProj {
  public:
    CBCTPolyForwardProjNoGrid(CTScanParas inCTScanParas, FilePath inFileInfo, const PhantomMaterial inPhantomMaterial);
    ~CBCTPolyForwardProjNoGrid();

    // Label volume -> per-material densities, all energy bins, with response.
    virtual void computePolyForwProj();
    // Same without detector response (the reference's run() left this
    // combination empty, ForwardProjection.cpp:60-64,72-75).
    virtual void computePolyForwProjNoResponse();
    // Flat field only, host side (no GPU).
    virtual void testI0();

  protected:
    const char *tag() const override { return "NoGrid"; }
};
