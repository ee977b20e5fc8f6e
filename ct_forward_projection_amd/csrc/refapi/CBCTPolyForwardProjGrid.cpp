#include "CBCTPolyForwardProjGrid.h"

// Constructor copies the scan, grid, path and material descriptions like the
// reference's (CBCTPolyForwardProjGrid.cpp:9-79); spectrumStep is 1.
CBCTPolyForwardProjGrid::CBCTPolyForwardProjGrid(CTScanParas inCTScanParas, GridInfo inGridInfo,
                                                 FilePath inFileInfo,
                                                 const PhantomMaterial inPhantomMaterial)
{
    mCTScanParas = inCTScanParas;
    mCTScanParas.spectrumStep = 1;
    mGridInfo = inGridInfo;
    mFilePath = inFileInfo;
    _mPhantomMaterial = inPhantomMaterial;
}

CBCTPolyForwardProjGrid::~CBCTPolyForwardProjGrid() {}

void CBCTPolyForwardProjGrid::computePolyForwProj() { runScan(false, true, true); }
void CBCTPolyForwardProjGrid::computePolyForwProjNoResponse() { runScan(false, true, false); }
void CBCTPolyForwardProjGrid::testI0() { runTestI0(true); }
