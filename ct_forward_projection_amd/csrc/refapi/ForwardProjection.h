// ForwardProjection.h -- INI-driven driver with the interface of the
// reference's ForwardProjection (Source/ForwardProjection.h:11-65): run()
// reads Config/ScanSetting.ini and Config/OutPutFolder.ini relative to the
// working directory and dispatches on IsTestI0 / IsDetResponse / IsGrid /
// IsSingleMaterial.  Invalid configuration throws ctfp::Error(CTFP_ERR_CONFIG).
#pragma once
#include <string>

#include "CBCTPolyForwardProjGrid.h"
#include "CBCTPolyForwardProjNoGrid.h"
#include "CBCTSinMatPolyForwardProjGrid.h"
#include "CBCTSinMatPolyForwardProjNoGrid.h"
#include "DataType.h"
#include "InitFile.h"

class ForwardProjection {
  public:
    ForwardProjection();
    ~ForwardProjection();

    void run();

    // Directory holding Config/ (default: the working directory).
    std::string rootDir = ".";
    int device = 0;

  private:
    void forwardPolyProjGrid();
    void forwardPolyProjGridNoResponse();
    void forwardSinMatPolyProjGrid();
    void forwardSinMatPolyProjGridNoResponse();
    void forwardPolyProjNoGrid();
    void forwardPolyProjNoGridNoResponse();
    void forwardSinMatPolyProjNoGrid();
    void forwardSinMatPolyProjNoGridNoResponse();
    void forwarProjNoGridTestI0();

    void _getCTScanConfig();
    void _getCTScanGridConfig();
    void _getCTScanDetResponseConfig();
    void _getOutputFolderConfig();
    std::string resolve(const std::string &p) const;

    CTScanParas mCTScanParas;
    GridInfo mGridInfo;
    FilePath mFilePath;
    CInitFile mCInitFile;
    int _isTestI0 = 0, _isDetResponse = 0, _isGrid = 0, _isSingleMaterial = 0;
    PhantomMaterial _mPhantomMaterial;
    float *_density = nullptr;
    std::string *_massAttenPath = nullptr;
};
