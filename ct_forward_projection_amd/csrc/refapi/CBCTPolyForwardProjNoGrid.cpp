#include "CBCTPolyForwardProjNoGrid.h"

// CBCTPolyForwardProjNoGrid.cpp:11-67; spectrumStep is 1.
CBCTPolyForwardProjNoGrid::CBCTPolyForwardProjNoGrid(CTScanParas inCTScanParas, FilePath inFileInfo,
                                                     const PhantomMaterial inPhantomMaterial)
{
    mCTScanParas = inCTScanParas;
    mCTScanParas.spectrumStep = 1;
    mFilePath = inFileInfo;
    _mPhantomMaterial = inPhantomMaterial;
}

CBCTPolyForwardProjNoGrid::~CBCTPolyForwardProjNoGrid() {}

void CBCTPolyForwardProjNoGrid::computePolyForwProj() { runScan(false, false, true); }
void CBCTPolyForwardProjNoGrid::computePolyForwProjNoResponse() { runScan(false, false, false); }
void CBCTPolyForwardProjNoGrid::testI0() { runTestI0(false); }
