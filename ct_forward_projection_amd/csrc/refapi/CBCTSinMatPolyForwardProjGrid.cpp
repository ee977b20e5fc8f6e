#include "CBCTSinMatPolyForwardProjGrid.h"

CBCTSinMatPolyForwardProjGrid::~CBCTSinMatPolyForwardProjGrid() {}
// CBCTSinMatPolyForwardProjGrid.cpp:13-145 / 147-267
void CBCTSinMatPolyForwardProjGrid::computePolyForwProj() { runScan(true, true, true); }
void CBCTSinMatPolyForwardProjGrid::computePolyForwProjNoResponse() { runScan(true, true, false); }
void CBCTSinMatPolyForwardProjGrid::computePolyForwProjFoSp() { runScan(true, true, true, true); }
