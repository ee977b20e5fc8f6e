#include "CBCTForwardProj.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>

#include "ctfp.h"

namespace ctfp {

std::string nativePath(const std::string &p)
{
    std::string s = p;
    for (char &c : s)
        if (c == '\\') c = '/';
    return s;
}

void readTable(const std::string &path, int n, float *out, const char *what)
{
    std::ifstream f(nativePath(path));
    if (!f.is_open()) throw Error(CTFP_ERR_IO, std::string("cannot open ") + what + " table " + path);
    int i = 0;
    std::string tok;
    while (i < n && (f >> tok)) {
        // every value must be a finite number: a NaN or an infinity would
        // reach every ray through the energy sums
        char *end = nullptr;
        const float v = std::strtof(tok.c_str(), &end);
        if (end == tok.c_str() || *end != '\0' || !std::isfinite(v))
            throw Error(CTFP_ERR_INVALID, std::string(what) + " table " + path + ": value " + std::to_string(i + 1) +
                                              " (\"" + tok + "\") is not a finite number");
        out[i++] = v;
    }
    if (i < n)
        throw Error(CTFP_ERR_IO, std::string(what) + " table " + path + " has " + std::to_string(i) +
                                     " values, " + std::to_string(n) + " needed");
}

} // namespace ctfp

namespace {
void check(int rc)
{
    if (rc != CTFP_OK) throw ctfp::Error(rc, ctfp_last_error());
}

struct Ctx {
    ctfp_context *c = nullptr;
    ~Ctx() { ctfp_destroy(c); }
};
} // namespace

CBCTForwardProj::CBCTForwardProj() {}
CBCTForwardProj::~CBCTForwardProj() { delete mPhantomObject; }

void CBCTForwardProj::computeParas()
{
    ctfp_scan_params s{};
    s.p_num_x = (int64_t)mCTScanParas.pNumX;
    s.p_num_y = (int64_t)mCTScanParas.pNumY;
    s.p_num_z = (int64_t)mCTScanParas.pNumZ;
    s.p_length_x = mCTScanParas.pLengthX;
    s.p_length_y = mCTScanParas.pLengthY;
    s.p_length_z = mCTScanParas.pLengthZ;
    s.proj_num = (int64_t)mCTScanParas.projNum;
    s.d_num_u = (int64_t)mCTScanParas.dNumU;
    s.d_num_v = (int64_t)mCTScanParas.dNumV;
    s.sod = mCTScanParas.sod;
    s.sdd = mCTScanParas.sdd;
    s.rotated_direction = mCTScanParas.rotatedDirection;
    s.d_size = mCTScanParas.dSize;
    ctfp_system_info o;
    check(ctfp_compute_paras(&s, &o));
    CTScanSystemInfo &y = mCTScanSystemInfo;
    y.rotatedDirection = o.rotated_direction;
    y.thetaStep = o.theta_step;
    y.dHalfLU = o.d_half_lu;
    y.dHalfLV = o.d_half_lv;
    y.FOVR = o.fov_r;
    y.FOVH = o.fov_h;
    y.pHalfX = o.p_half_x;
    y.pHalfY = o.p_half_y;
    y.pHalfZ = o.p_half_z;
    y.pSizeX = o.p_size_x;
    y.pSizeY = o.p_size_y;
    y.pSizeZ = o.p_size_z;
    y.dx = o.dx;
    y.dy = o.dy;
    y.dz = o.dz;
    y.intNum = (size_t)o.int_num;
}

void CBCTForwardProj::initGridInfo()
{
    // CBCTPolyForwardProjGrid::initGridInfo (CBCTPolyForwardProjGrid.cpp:287-308)
    GridAndDetectorSystem &g = mGridAndDetectorSystem;
    g.angleNums = mCTScanParas.projNum;
    g.dNums = mCTScanParas.dNumU;
    g.dSize = mCTScanParas.dSize * 1000; // mm -> um, FP32
    g.FD = mGridInfo.FD;
    g.h = mGridInfo.h;
    g.gridStripsWidth = mGridInfo.gridStripsWidth;
    g.interspcaceWidth = mGridInfo.interspcaceWidth;
    g.gridStripMaterial = mGridInfo.gridStripMaterial;
    g.interspaceMaterial = mGridInfo.interspaceMaterial;
    g.gridStripDensity = mGridInfo.gridStripDensity;
    g.interspaceDensity = mGridInfo.interspaceDensity;
    g.gridStripMAPath = mGridInfo.gridStripMAPath;
    g.interspaceMAPath = mGridInfo.interspaceMAPath;
    g.mI0Photons = mCTScanParas.I0Val;
}

void CBCTForwardProj::allocHost(bool withGrid)
{
    const size_t npix = mCTScanParas.dNumU * mCTScanParas.dNumV;
    const size_t n = npix * mCTScanParas.projNum;
    const int E = mCTScanParas.specEnergyNum;
    vI.assign(n, 0.0f);
    vIA.assign(n, 0.0f);
    vI0.assign(npix, 0.0f);
    vSpec.assign(E, 0.0f);
    vDet.assign(npix, 0.0f);
    vADet.assign(E, 0.0f);
    I = vI.data();
    IAbsorb = vIA.data();
    // per-bin intensities (addIandIAbsorb): none, the device sums the bins
    h_mForwardProj.I = nullptr;
    h_mForwardProj.IAbsorb = nullptr;
    h_mForwardProj.I0 = vI0.data();
    h_mForwardProj.spectrumNormal = vSpec.data();
    h_mForwardProj.detResponse = vDet.data();
    aDetResponse = vADet.data();
    if (withGrid) {
        vGridRow.assign(mCTScanParas.dNumU, 1.0f);
        h_mForwardProj.grid = vGridRow.data();
    }
}

void CBCTForwardProj::readSpecrtumNorm()
{
    ctfp::readTable(mFilePath.spectrumPath, mCTScanParas.specEnergyNum, h_mForwardProj.spectrumNormal,
                    "spectrum");
}

void CBCTForwardProj::readGridMassAttu()
{
    // CBCTForwardProj.cpp:174-225: mass attenuation * density / 10 -> 1/mm.
    const int E = mCTScanParas.specEnergyNum;
    vGridMu.assign(E, 0.0f);
    vInterMu.assign(E, 0.0f);
    ctfp::readTable(mGridInfo.gridStripMAPath, E, vGridMu.data(), "grid strip");
    ctfp::readTable(mGridInfo.interspaceMAPath, E, vInterMu.data(), "grid interspace");
    // The reference's grid drivers never copy InterspaceDensity into their
    // GridInfo (CBCTPolyForwardProjGrid.cpp:53-63), so it is 0 there.
    const float rho_i = honorInterspaceDensity ? mGridInfo.interspaceDensity : 0.0f;
    for (int e = 0; e < E; e++) {
        vGridMu[e] = vGridMu[e] * mGridInfo.gridStripDensity / 10;
        vInterMu[e] = vInterMu[e] * rho_i / 10;
    }
    h_mForwardProj.gridLinearAtten = vGridMu.data();
    h_mForwardProj.interspaceLinearAtten = vInterMu.data();
}

void CBCTForwardProj::readScintillatorMassAttu()
{
    const int E = mCTScanParas.specEnergyNum;
    vScinMu.assign(E, 0.0f);
    ctfp::readTable(mFilePath.scintillatorPath, E, vScinMu.data(), "scintillator");
    for (int e = 0; e < E; e++)
        vScinMu[e] = vScinMu[e] * mCTScanParas.mScintilltorInfo.scintillatorDensity / 10;
    h_mForwardProj.scintillatorLineAtten = vScinMu.data();
}

void CBCTForwardProj::computePerScinThinckness()
{
    const size_t npix = mCTScanParas.dNumU * mCTScanParas.dNumV;
    vThick.assign(npix, 0.0f);
    check(ctfp_scint_thickness((int64_t)npix, mCTScanParas.mScintilltorInfo.scintillatorThickness,
                               mCTScanParas.mScintilltorInfo.scintillatorThicknessErr, vThick.data()));
    h_mForwardProj.scintillatorPerThickness = vThick.data();
}

void CBCTForwardProj::updateDetResponse()
{
    const size_t npix = mCTScanParas.dNumU * mCTScanParas.dNumV;
    check(ctfp_det_response((int64_t)npix, mCTScanParas.mScintilltorInfo.detResponseFactor,
                            h_mForwardProj.scintillatorLineAtten[specIndex],
                            h_mForwardProj.scintillatorPerThickness, h_mForwardProj.detResponse));
}

void CBCTForwardProj::getADetResponse()
{
    // CBCTForwardProj.cpp:257: flat index (NU + NV) / 2 (row 0, Q14)
    aDetResponse[specIndex] = h_mForwardProj.detResponse[(mCTScanParas.dNumU + mCTScanParas.dNumV) / 2];
}

void CBCTForwardProj::saveADetResponse()
{
    std::filesystem::create_directories(ctfp::nativePath(mFilePath.outputFolder));
    std::string path = ctfp::nativePath(mFilePath.outputFolder) + "/ADetResponse.txt";
    std::ofstream f(path);
    if (!f.is_open()) throw ctfp::Error(CTFP_ERR_IO, "cannot write " + path);
    for (int e = 0; e < mCTScanParas.specEnergyNum; e++) f << aDetResponse[e] << "\n";
}

// CBCTForwardProj.cpp:65-76: adds one energy bin's intensities
// (h_mForwardProj.I / IAbsorb, [projNum][dNumV][dNumU]) to the scan totals
// I / IAbsorb.  For callers that drive the per-energy loop themselves; the
// scans of this class sum every bin on the device and leave the per-bin
// pointers null, and then this adds nothing.
void CBCTForwardProj::addIandIAbsorb()
{
    const size_t n = vI.size();
    if (h_mForwardProj.I && h_mForwardProj.I != I)
        for (size_t i = 0; i < n; i++) I[i] += h_mForwardProj.I[i];
    if (h_mForwardProj.IAbsorb && h_mForwardProj.IAbsorb != IAbsorb)
        for (size_t i = 0; i < n; i++) IAbsorb[i] += h_mForwardProj.IAbsorb[i];
}

void CBCTForwardProj::addStandardI0(float &I0Standard)
{
    // CBCTForwardProj.cpp:92-95
    I0Standard += mCTScanParas.I0Val * mCTScanSystemInfo.spectrumVal *
                  mCTScanParas.mScintilltorInfo.detResponseFactor *
                  (1 - expf(-h_mForwardProj.scintillatorLineAtten[specIndex] *
                            mCTScanParas.mScintilltorInfo.scintillatorThickness));
}

void CBCTForwardProj::addI0()
{
    // addI0 / addI0NoResponse (CBCTPolyForwardProjGrid.cpp:330-350,
    // CBCTPolyForwardProjNoGrid.cpp:290-310) for the current bin.
    const size_t nu = mCTScanParas.dNumU, nv = mCTScanParas.dNumV;
    for (size_t v = 0; v < nv; v++)
        for (size_t u = 0; u < nu; u++) {
            float w = mCTScanParas.I0Val * mCTScanSystemInfo.spectrumVal;
            if (hasGrid_) w = w * h_mForwardProj.grid[u];
            if (hasResponse_) w = w * h_mForwardProj.detResponse[v * nu + u];
            h_mForwardProj.I0[v * nu + u] += w;
        }
}

void CBCTForwardProj::convertIandIAbsorbDt()
{
    const size_t n = vI.size();
    vIu.resize(n);
    vIAu.resize(n);
    for (size_t i = 0; i < n; i++) {
        vIu[i] = (uint32)floorf(I[i]);
        vIAu[i] = (uint32)floorf(IAbsorb[i]);
    }
    IPolyenergetic = vIu.data();
    IPolyenergeticAbsorb = vIAu.data();
}

void CBCTForwardProj::computeProj()
{
    const size_t npix = mCTScanParas.dNumU * mCTScanParas.dNumV;
    vProj.assign(vI.size(), 0.0f);
    for (size_t j = 0; j < mCTScanParas.projNum; j++)
        for (size_t i = 0; i < npix; i++)
            vProj[j * npix + i] = logf(h_mForwardProj.I0[i] / (float)IPolyenergetic[j * npix + i]);
    h_mForwardProj.proj = vProj.data();
}

namespace {
void writeRaw(const std::string &path, const void *data, size_t bytes)
{
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw ctfp::Error(CTFP_ERR_IO, "cannot write " + path);
    f.write((const char *)data, (std::streamsize)bytes);
    if (!f) throw ctfp::Error(CTFP_ERR_IO, "short write to " + path);
}
} // namespace

void CBCTForwardProj::saveIandIAbsorb()
{
    // CBCTPolyForwardProjGrid.cpp:310-327 / CBCTPolyForwardProjNoGrid.cpp:271-288
    std::string dims = std::to_string(mCTScanParas.dNumU) + "x" + std::to_string(mCTScanParas.dNumV) +
                       "x" + std::to_string(mCTScanParas.projNum);
    std::string dir = ctfp::nativePath(mFilePath.outputFolder);
    writeRaw(dir + "/PhoImgPro_" + tag() + "_" + dims + "_uint32.raw", IPolyenergetic,
             vIu.size() * sizeof(uint32));
    writeRaw(dir + "/PhoImgAbsorb_" + tag() + "_" + dims + "_uint32.raw", IPolyenergeticAbsorb,
             vIAu.size() * sizeof(uint32));
}

void CBCTForwardProj::saveI0()
{
    std::string dir = ctfp::nativePath(mFilePath.outputFolder);
    writeRaw(dir + "/BgImgPro_" + tag() + "_" + std::to_string(mCTScanParas.dNumU) + "x" +
                 std::to_string(mCTScanParas.dNumV) + "_float.raw",
             h_mForwardProj.I0, vI0.size() * sizeof(float));
}

void CBCTForwardProj::saveProj()
{
    std::string dir = ctfp::nativePath(mFilePath.outputFolder);
    writeRaw(dir + "/Projection_" + tag() + "_" + std::to_string(mCTScanParas.dNumU) + "x" +
                 std::to_string(mCTScanParas.dNumV) + "x" + std::to_string(mCTScanParas.projNum) +
                 "_float.raw",
             h_mForwardProj.proj, vProj.size() * sizeof(float));
}

void CBCTForwardProj::creatOutputFolder()
{
    std::filesystem::path p(ctfp::nativePath(mFilePath.outputFolder));
    if (cleanOutputFolder && std::filesystem::exists(p)) _removeOutputFolder();
    std::error_code ec;
    std::filesystem::create_directories(p, ec);
    if (ec) throw ctfp::Error(CTFP_ERR_IO, "cannot create " + p.string() + ": " + ec.message());
}

void CBCTForwardProj::_removeOutputFolder()
{
    std::error_code ec;
    std::filesystem::remove_all(ctfp::nativePath(mFilePath.outputFolder), ec);
    if (ec) throw ctfp::Error(CTFP_ERR_IO, "cannot remove " + mFilePath.outputFolder);
}

void CBCTForwardProj::showScanSystemInfo()
{
    if (!verbose) return;
    std::cout << "******************* scan system *******************\n"
              << "detector       : " << mCTScanParas.dNumU << " x " << mCTScanParas.dNumV << " @ "
              << mCTScanParas.dSize << " mm\n"
              << "SDD / SOD      : " << mCTScanParas.sdd << " / " << mCTScanParas.sod << " mm\n"
              << "FOV (d x h)    : " << 2 * mCTScanSystemInfo.FOVR << " x " << 2 * mCTScanSystemInfo.FOVH << " mm\n"
              << "voxel          : " << mCTScanSystemInfo.pSizeX << " x " << mCTScanSystemInfo.pSizeY << " x "
              << mCTScanSystemInfo.pSizeZ << " mm\n"
              << "samples / ray  : " << mCTScanSystemInfo.intNum << "\n"
              << "photons (I0Val): " << mCTScanParas.I0Val << "\n\n";
}

void CBCTForwardProj::showGridmInfo()
{
    if (!verbose) return;
    std::cout << "******************* focused grid ******************\n"
              << "strip          : " << mGridInfo.gridStripMaterial << ", " << mGridInfo.gridStripsWidth << " um\n"
              << "interspace     : " << mGridInfo.interspaceMaterial << ", " << mGridInfo.interspcaceWidth << " um\n"
              << "period         : " << mGridInfo.gridStripsWidth + mGridInfo.interspcaceWidth << " um\n"
              << "height / FD    : " << mGridInfo.h << " / " << mGridInfo.FD << " mm\n\n";
}

void CBCTForwardProj::showProcessInfo()
{
    if (!verbose) return;
    std::cout << (specIndex + 1) * mCTScanParas.spectrumStep << " keV: S = "
              << h_mForwardProj.spectrumNormal[specIndex] << "\n";
}

std::vector<float> CBCTForwardProj::gridTable()
{
    const int E = mCTScanParas.specEnergyNum;
    std::vector<float> T((size_t)E * mCTScanParas.dNumU);
    Grid g(h_mForwardProj.grid, mGridAndDetectorSystem);
    g.defGrid();
    g.transmission(E, vGridMu.data(), vInterMu.data(), T.data());
    // the last bin stays bound as the current row, like the reference's loop
    std::copy(T.end() - mCTScanParas.dNumU, T.end(), h_mForwardProj.grid);
    g.getPeriodIncex(gridPeriodIndex);
    return T;
}

// computeFoSpOffset (CBCTForwardProj.cpp:122-148): per-view focal-spot
// offsets within a disc of diameter focalSpotSize (mm), from MSVC rand()
// after srand(10) as in the reference, but kept inside the disc.
void CBCTForwardProj::computeFoSpOffset()
{
    const size_t n = mCTScanParas.projNum;
    vFoSpU.assign(n, 0.0f);
    vFoSpV.assign(n, 0.0f);
    check(ctfp_focal_spot_offsets((int64_t)n, mCTScanParas.focalSpotSize, 10u, vFoSpU.data(), vFoSpV.data()));
    h_mForwardProj.foSpOffsetU = vFoSpU.data();
    h_mForwardProj.foSpOffsetV = vFoSpV.data();
}

void CBCTForwardProj::runScan(bool singleMaterial, bool withGrid, bool withResponse, bool focalSpot)
{
    hasGrid_ = withGrid;
    hasResponse_ = withResponse;
    allocHost(withGrid);
    computeParas();
    showScanSystemInfo();
    if (withGrid) showGridmInfo();
    readSpecrtumNorm();
    delete mPhantomObject;
    mPhantomObject = new PhantomObject(mCTScanParas, _mPhantomMaterial);
    mPhantomObject->readData();
    const int E = mCTScanParas.specEnergyNum;
    const int M = singleMaterial ? 1 : _mPhantomMaterial.materialNum;
    if (withResponse) {
        computePerScinThinckness();
        readScintillatorMassAttu();
    }
    if (withGrid) {
        readGridMassAttu();
        initGridInfo();
        vGridT = gridTable();
    }
    // attenuation per unit density, 1/mm per g/cm^3
    std::vector<float> atten((size_t)E * M);
    for (int e = 0; e < E; e++)
        for (int m = 0; m < M; m++) atten[(size_t)e * M + m] = mPhantomObject->massAtten(e, m) / 10.f;

    ctfp_scan_params s{};
    s.p_num_x = (int64_t)mCTScanParas.pNumX;
    s.p_num_y = (int64_t)mCTScanParas.pNumY;
    s.p_num_z = (int64_t)mCTScanParas.pNumZ;
    s.p_length_x = mCTScanParas.pLengthX;
    s.p_length_y = mCTScanParas.pLengthY;
    s.p_length_z = mCTScanParas.pLengthZ;
    s.proj_num = (int64_t)mCTScanParas.projNum;
    s.d_num_u = (int64_t)mCTScanParas.dNumU;
    s.d_num_v = (int64_t)mCTScanParas.dNumV;
    s.sod = mCTScanParas.sod;
    s.sdd = mCTScanParas.sdd;
    s.rotated_direction = mCTScanParas.rotatedDirection;
    s.d_size = mCTScanParas.dSize;
    Ctx ctx;
    check(ctfp_create(&s, device, &ctx.c));
    check(ctfp_set_labels(ctx.c, mPhantomObject->labels(), 0, singleMaterial ? 1 : M,
                          mPhantomObject->densities(), singleMaterial ? 1 : 0, CTFP_F32));
    ctfp_spectral sp{};
    sp.n_energy = E;
    sp.n_mat = M;
    sp.spectrum = h_mForwardProj.spectrumNormal;
    sp.atten = atten.data();
    sp.i0_val = (float)mCTScanParas.I0Val;
    sp.grid = withGrid ? vGridT.data() : nullptr;
    if (withResponse) {
        sp.scint_mu = vScinMu.data();
        sp.scint_thickness = vThick.data();
        sp.det_factor = mCTScanParas.mScintilltorInfo.detResponseFactor;
    }
    check(ctfp_set_spectral(ctx.c, &sp));
    if (focalSpot) {
        computeFoSpOffset();
        check(ctfp_set_focal_spot(ctx.c, h_mForwardProj.foSpOffsetU, h_mForwardProj.foSpOffsetV));
    }
    if (verbose) std::cout << "projecting " << mCTScanParas.projNum << " views, " << E << " energy bins\n";
    // the reference times each GPU call, copies to the host included
    // (CBCTPolyProjGPU.cu:94-101)
    const auto t0 = std::chrono::steady_clock::now();
    check(ctfp_forward_host(ctx.c, I, IAbsorb, h_mForwardProj.I0));
    mCTScanSystemInfo.totalTime = std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count();
    if (verbose) std::cout << "==>> projection time (GPU, with the copy to the host): " << mCTScanSystemInfo.totalTime
                           << " s\n";

    if (withResponse) { // response at the reference's "middle" pixel per bin
        for (specIndex = 0; specIndex < (size_t)E; specIndex++) {
            updateDetResponse();
            getADetResponse();
        }
        specIndex = 0;
    }
    convertIandIAbsorbDt();
    creatOutputFolder();
    saveIandIAbsorb();
    saveI0();
    if (withGrid) {
        Grid g(h_mForwardProj.grid, mGridAndDetectorSystem);
        g.getPeriodIncex(gridPeriodIndex);
        g.savePeriodIndex(ctfp::nativePath(mFilePath.outputFolder) + "/GridPeriodIndex_" +
                          std::to_string(mCTScanParas.dNumU) + ".txt");
    }
    if (singleMaterial && withResponse) saveADetResponse();
    computeProj();
    saveProj();
    if (verbose) std::cout << "forward projection done -> " << mFilePath.outputFolder << "\n";
}

void CBCTForwardProj::runTestI0(bool withGrid)
{
    // testI0 (CBCTPolyForwardProjNoGrid.cpp:188-264, CBCTPolyForwardProjGrid.cpp:205-284):
    // host only, flat field of the response (and grid) in bin order.
    hasGrid_ = withGrid;
    hasResponse_ = true;
    allocHost(withGrid);
    computeParas();
    readSpecrtumNorm();
    computePerScinThinckness();
    readScintillatorMassAttu();
    if (withGrid) {
        readGridMassAttu();
        initGridInfo();
        vGridT = gridTable();
    }
    float I0Standard = 0;
    for (specIndex = 0; specIndex < (size_t)mCTScanParas.specEnergyNum; specIndex++) {
        mCTScanSystemInfo.spectrumVal = h_mForwardProj.spectrumNormal[specIndex];
        updateDetResponse();
        if (withGrid)
            std::copy(vGridT.begin() + specIndex * mCTScanParas.dNumU,
                      vGridT.begin() + (specIndex + 1) * mCTScanParas.dNumU, h_mForwardProj.grid);
        addI0();
        addStandardI0(I0Standard);
    }
    specIndex = 0;
    creatOutputFolder();
    saveI0();
    if (verbose)
        std::cout << "I0 test done: nominal scintillator thickness "
                  << mCTScanParas.mScintilltorInfo.scintillatorThickness << " mm gives I0 = " << I0Standard
                  << " photons\n";
    lastI0Standard = I0Standard;
}
