#include "ForwardProjection.h"

#include <filesystem>
#include <iostream>
#include <strings.h>

#include "ctfp.h"

namespace {
[[noreturn]] void bad(const std::string &what)
{
    throw ctfp::Error(CTFP_ERR_CONFIG, "configuration error: " + what);
}
} // namespace

ForwardProjection::ForwardProjection() {}

ForwardProjection::~ForwardProjection()
{
    delete[] _density;
    delete[] _massAttenPath;
}

std::string ForwardProjection::resolve(const std::string &p) const
{
    std::string n = ctfp::nativePath(p);
    if (n.empty() || n[0] == '/' || rootDir == ".") return n;
    return rootDir + "/" + n;
}

void ForwardProjection::run()
{
    _getCTScanConfig();
    _getOutputFolderConfig();
    // Dispatch of ForwardProjection.cpp:25-83.  The two no-response
    // multi-material branches were empty in the reference; they run here.
    if (_isTestI0) return forwarProjNoGridTestI0();
    if (_isDetResponse) {
        if (_isGrid) return _isSingleMaterial ? forwardSinMatPolyProjGrid() : forwardPolyProjGrid();
        return _isSingleMaterial ? forwardSinMatPolyProjNoGrid() : forwardPolyProjNoGrid();
    }
    if (_isGrid)
        return _isSingleMaterial ? forwardSinMatPolyProjGridNoResponse() : forwardPolyProjGridNoResponse();
    return _isSingleMaterial ? forwardSinMatPolyProjNoGridNoResponse() : forwardPolyProjNoGridNoResponse();
}

void ForwardProjection::_getCTScanConfig()
{
    // ForwardProjection.cpp:197-323, sentinel -100 / "null" for missing keys
    std::string file = resolve("Config/ScanSetting.ini");
    mCInitFile.GetFileName(file.c_str());
    if (!mCInitFile.IsLoaded()) bad("cannot read " + file);
    CInitFile &F = mCInitFile;
    const char *sec = "ScanPara";
    F.GetEntryValue(sec, "IsTestI0", -100, _isTestI0);
    F.GetEntryValue(sec, "IsDetResponse", -100, _isDetResponse);
    F.GetEntryValue(sec, "IsGrid", -100, _isGrid);
    F.GetEntryValue(sec, "IsSingleMaterial", -100, _isSingleMaterial);
    F.GetEntryValue(sec, "SDD", -100.0f, mCTScanParas.sdd);
    F.GetEntryValue(sec, "SOD", -100.0f, mCTScanParas.sod);
    bool okV = F.GetEntryValue(sec, "ScanAngleNum", (size_t)0, mCTScanParas.projNum);
    F.GetEntryValue(sec, "RotatedDirection", -100, mCTScanParas.rotatedDirection);
    // optional, not read by the reference (which fixes 0.4 mm): used only by
    // the focal-spot projections (computePolyForwProjFoSp)
    float spot = -100.0f;
    F.GetEntryValue(sec, "FocalSpotSize", -100.0f, spot);
    if (spot != -100.0f) {
        if (!(spot > 0.0f) || spot > 100.0f) bad("[ScanPara] FocalSpotSize must be in (0, 100] mm");
        mCTScanParas.focalSpotSize = spot;
    }
    for (int flag : {_isTestI0, _isDetResponse, _isGrid, _isSingleMaterial})
        if (flag != 0 && flag != 1) bad("[ScanPara] IsTestI0/IsDetResponse/IsGrid/IsSingleMaterial must be 0 or 1");
    if (!(mCTScanParas.sdd > 0) || !(mCTScanParas.sod > 0) || !(mCTScanParas.sdd > mCTScanParas.sod))
        bad("[ScanPara] SDD > SOD > 0 required");
    if (!okV || mCTScanParas.projNum == 0) bad("[ScanPara] ScanAngleNum must be a positive integer");
    if (mCTScanParas.rotatedDirection != 1 && mCTScanParas.rotatedDirection != -1)
        bad("[ScanPara] RotatedDirection must be 1 or -1");

    sec = "RaySource";
    bool okP = F.GetEntryValue(sec, "Photons", (size_t)0, mCTScanParas.I0Val);
    F.GetEntryValue(sec, "SpecEnergyNum", -100, mCTScanParas.specEnergyNum);
    F.GetEntryValue(sec, "SpectrumPath", "null", mFilePath.spectrumPath);
    if (!okP || mCTScanParas.specEnergyNum < 1 || mFilePath.spectrumPath == "null")
        bad("[RaySource] Photons, SpecEnergyNum and SpectrumPath are required");
    mFilePath.spectrumPath = resolve(mFilePath.spectrumPath);

    sec = "PhantomVox";
    bool ok = F.GetEntryValue(sec, "PNumX", (size_t)0, mCTScanParas.pNumX);
    ok = F.GetEntryValue(sec, "PNumY", (size_t)0, mCTScanParas.pNumY) && ok;
    ok = F.GetEntryValue(sec, "PNumZ", (size_t)0, mCTScanParas.pNumZ) && ok;
    if (!ok || !mCTScanParas.pNumX || !mCTScanParas.pNumY || !mCTScanParas.pNumZ)
        bad("[PhantomVox] PNumX/Y/Z must be positive integers");
    F.GetEntryValue(sec, "PLengthX", -100.0f, mCTScanParas.pLengthX);
    F.GetEntryValue(sec, "PLengthY", -100.0f, mCTScanParas.pLengthY);
    F.GetEntryValue(sec, "PLengthZ", -100.0f, mCTScanParas.pLengthZ);
    if (!(mCTScanParas.pLengthX > 0) || !(mCTScanParas.pLengthY > 0) || !(mCTScanParas.pLengthZ > 0))
        bad("[PhantomVox] PLengthX/Y/Z must be positive");

    sec = "Detector";
    ok = F.GetEntryValue(sec, "DNumU", (size_t)0, mCTScanParas.dNumU);
    ok = F.GetEntryValue(sec, "DNumV", (size_t)0, mCTScanParas.dNumV) && ok;
    F.GetEntryValue(sec, "DSize", -100.0f, mCTScanParas.dSize);
    if (!ok || !mCTScanParas.dNumU || !mCTScanParas.dNumV || !(mCTScanParas.dSize > 0))
        bad("[Detector] DNumU, DNumV and DSize must be positive");

    sec = "PhantomMaterial";
    F.GetEntryValue(sec, "MaterialNum", -100, _mPhantomMaterial.materialNum);
    F.GetEntryValue(sec, "StructureMarkPath", "null", _mPhantomMaterial.structureMarkPath);
    if (_mPhantomMaterial.materialNum < 1 || _mPhantomMaterial.materialNum > CTFP_MAX_MATERIALS ||
        _mPhantomMaterial.structureMarkPath == "null")
        bad("[PhantomMaterial] MaterialNum (1.." + std::to_string(CTFP_MAX_MATERIALS) +
            ") and StructureMarkPath are required");
    _mPhantomMaterial.structureMarkPath = resolve(_mPhantomMaterial.structureMarkPath);
    const int M = _mPhantomMaterial.materialNum;
    delete[] _density;
    delete[] _massAttenPath;
    _density = new float[M];
    _massAttenPath = new std::string[M];
    for (int i = 0; i < M; i++) {
        std::string dk = "Density" + std::to_string(i + 1), mk = "MassAttenPath" + std::to_string(i + 1);
        F.GetEntryValue(sec, dk.c_str(), -100.0f, _density[i]);
        F.GetEntryValue(sec, mk.c_str(), "null", _massAttenPath[i]);
        if (!(_density[i] >= 0) || _massAttenPath[i] == "null")
            bad("[PhantomMaterial] " + dk + " / " + mk + " missing or invalid");
        _massAttenPath[i] = resolve(_massAttenPath[i]);
    }
    _mPhantomMaterial.density = _density;
    _mPhantomMaterial.massAttenPath = _massAttenPath;
}

void ForwardProjection::_getCTScanGridConfig()
{
    // ForwardProjection.cpp:326-348
    CInitFile &F = mCInitFile;
    const char *sec = "PostGrid";
    F.GetEntryValue(sec, "FD", -100.0f, mGridInfo.FD);
    F.GetEntryValue(sec, "Height", -100.0f, mGridInfo.h);
    F.GetEntryValue(sec, "GridStripsWidth", -100, mGridInfo.gridStripsWidth);
    F.GetEntryValue(sec, "InterspcaceWidth", -100, mGridInfo.interspcaceWidth);
    F.GetEntryValue(sec, "GridStripMaterial", "null", mGridInfo.gridStripMaterial);
    F.GetEntryValue(sec, "InterspaceMaterial", "null", mGridInfo.interspaceMaterial);
    F.GetEntryValue(sec, "GridStripDensity", -100.0f, mGridInfo.gridStripDensity);
    F.GetEntryValue(sec, "InterspaceDensity", -100.0f, mGridInfo.interspaceDensity);
    F.GetEntryValue(sec, "GridStripMAPath", "null", mGridInfo.gridStripMAPath);
    F.GetEntryValue(sec, "InterspaceMAPath", "null", mGridInfo.interspaceMAPath);
    if (!(mGridInfo.FD > 0) || !(mGridInfo.h >= 0) || mGridInfo.gridStripsWidth < 0 ||
        mGridInfo.interspcaceWidth < 1 || mGridInfo.interspaceMaterial == "null" ||
        mGridInfo.gridStripMaterial == "null" || !(mGridInfo.gridStripDensity >= 0) ||
        !(mGridInfo.interspaceDensity >= 0) || mGridInfo.gridStripMAPath == "null" ||
        mGridInfo.interspaceMAPath == "null")
        bad("[PostGrid] incomplete or invalid grid description");
    mGridInfo.gridStripMAPath = resolve(mGridInfo.gridStripMAPath);
    mGridInfo.interspaceMAPath = resolve(mGridInfo.interspaceMAPath);
}

void ForwardProjection::_getCTScanDetResponseConfig()
{
    // ForwardProjection.cpp:352-369
    CInitFile &F = mCInitFile;
    ScintilltorInfo &s = mCTScanParas.mScintilltorInfo;
    const char *sec = "Scintillator";
    F.GetEntryValue(sec, "Density", -100.0f, s.scintillatorDensity);
    F.GetEntryValue(sec, "Thickness", -100.0f, s.scintillatorThickness);
    F.GetEntryValue(sec, "ThicknessErr", -100.0f, s.scintillatorThicknessErr);
    F.GetEntryValue(sec, "PhotoelectConverFactor", -100.0f, s.detResponseFactor);
    F.GetEntryValue(sec, "ScintillatorMAPath", "null", mFilePath.scintillatorPath);
    if (!(s.scintillatorDensity >= 0) || !(s.scintillatorThickness >= 0) ||
        !(s.scintillatorThicknessErr >= 0) || !(s.detResponseFactor >= 0) || s.detResponseFactor > 1 ||
        mFilePath.scintillatorPath == "null")
        bad("[Scintillator] incomplete or invalid scintillator description");
    mFilePath.scintillatorPath = resolve(mFilePath.scintillatorPath);
}

void ForwardProjection::_getOutputFolderConfig()
{
    // ForwardProjection.cpp:372-419.  The reference opens Config/OutputFolder.ini
    // while the file is Config/OutPutFolder.ini (Q18): the name is matched
    // case-insensitively.
    std::string cfg = resolve("Config");
    std::string file;
    std::error_code ec;
    for (auto &e : std::filesystem::directory_iterator(cfg, ec))
        if (strcasecmp(e.path().filename().string().c_str(), "OutputFolder.ini") == 0)
            file = e.path().string();
    if (file.empty()) bad("cannot find " + cfg + "/OutPutFolder.ini");
    CInitFile F;
    F.GetFileName(file.c_str());
    const char *sec = _isTestI0 ? "NoGridResponseTestI0"
                      : _isGrid ? (_isDetResponse ? "GridResponse" : "GridNoResponse")
                                : (_isDetResponse ? "NoGridResponse" : "NoGridNoResponse");
    F.GetEntryValue(sec, "OutFolder", "null", mFilePath.outputFolder);
    if (mFilePath.outputFolder == "null") bad(std::string("[") + sec + "] OutFolder is required");
    mFilePath.outputFolder = resolve(mFilePath.outputFolder);
}

#define CTFP_RUN(Class, method, ...)        \
    {                                        \
        Class p(__VA_ARGS__);                \
        p.device = device;                   \
        p.method();                          \
    }

void ForwardProjection::forwardPolyProjGrid()
{
    _getCTScanGridConfig();
    _getCTScanDetResponseConfig();
    CTFP_RUN(CBCTPolyForwardProjGrid, computePolyForwProj, mCTScanParas, mGridInfo, mFilePath, _mPhantomMaterial)
}
void ForwardProjection::forwardPolyProjGridNoResponse()
{
    _getCTScanGridConfig();
    CTFP_RUN(CBCTPolyForwardProjGrid, computePolyForwProjNoResponse, mCTScanParas, mGridInfo, mFilePath,
             _mPhantomMaterial)
}
void ForwardProjection::forwardSinMatPolyProjGrid()
{
    _getCTScanGridConfig();
    _getCTScanDetResponseConfig();
    CTFP_RUN(CBCTSinMatPolyForwardProjGrid, computePolyForwProj, mCTScanParas, mGridInfo, mFilePath,
             _mPhantomMaterial)
}
void ForwardProjection::forwardSinMatPolyProjGridNoResponse()
{
    _getCTScanGridConfig();
    CTFP_RUN(CBCTSinMatPolyForwardProjGrid, computePolyForwProjNoResponse, mCTScanParas, mGridInfo, mFilePath,
             _mPhantomMaterial)
}
void ForwardProjection::forwardPolyProjNoGrid()
{
    _getCTScanDetResponseConfig();
    CTFP_RUN(CBCTPolyForwardProjNoGrid, computePolyForwProj, mCTScanParas, mFilePath, _mPhantomMaterial)
}
void ForwardProjection::forwardPolyProjNoGridNoResponse()
{
    CTFP_RUN(CBCTPolyForwardProjNoGrid, computePolyForwProjNoResponse, mCTScanParas, mFilePath, _mPhantomMaterial)
}
void ForwardProjection::forwardSinMatPolyProjNoGrid()
{
    _getCTScanDetResponseConfig();
    CTFP_RUN(CBCTSinMatPolyForwardProjNoGrid, computePolyForwProj, mCTScanParas, mFilePath, _mPhantomMaterial)
}
void ForwardProjection::forwardSinMatPolyProjNoGridNoResponse()
{
    CTFP_RUN(CBCTSinMatPolyForwardProjNoGrid, computePolyForwProjNoResponse, mCTScanParas, mFilePath,
             _mPhantomMaterial)
}
void ForwardProjection::forwarProjNoGridTestI0()
{
    _getCTScanDetResponseConfig();
    CTFP_RUN(CBCTPolyForwardProjNoGrid, testI0, mCTScanParas, mFilePath, _mPhantomMaterial)
}
