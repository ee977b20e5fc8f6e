#include "CBCTSinMatPolyForwardProjNoGrid.h"

CBCTSinMatPolyForwardProjNoGrid::~CBCTSinMatPolyForwardProjNoGrid() {}
// CBCTSinMatPolyForwardProjNoGrid.cpp:10-111 / 113-214
void CBCTSinMatPolyForwardProjNoGrid::computePolyForwProj() { runScan(true, false, true); }
void CBCTSinMatPolyForwardProjNoGrid::computePolyForwProjNoResponse() { runScan(true, false, false); }
void CBCTSinMatPolyForwardProjNoGrid::computePolyForwProjFoSp() { runScan(true, false, true, true); }
