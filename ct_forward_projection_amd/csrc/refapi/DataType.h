// DataType.h -- parameter structs of the reference's class API
// (Source/DataType.h:23-180), kept member-for-member so code written against
// the reference compiles unchanged.  Host-side only; the device boundary is
// the POD C ABI in include/ctfp.h.
#pragma once
#include <cstddef>
#include <string>

typedef unsigned char uint8;
typedef unsigned int uint32;

#define DELETEARR(arrP)      \
    if (arrP != nullptr) {   \
        delete[] arrP;       \
        arrP = nullptr;      \
    }

struct ScintilltorInfo {
    std::string scintillatorName;
    float scintillatorDensity = 0;      // g/cm^3
    float scintillatorThickness = 0;    // mm
    float scintillatorThicknessErr = 0; // mm
    float detResponseFactor = 0;
};

struct CTScanParas {
    size_t pNumX = 0, pNumY = 0, pNumZ = 0;       // voxels
    float pLengthX = 0, pLengthY = 0, pLengthZ = 0; // mm
    size_t projNum = 0;
    size_t dNumU = 0, dNumV = 0;
    float sod = 0, sdd = 0; // mm
    int rotatedDirection = 1;
    float dSize = 0; // mm
    size_t I0Val = 0;
    float focalSpotSize = 0.4f; // mm; the reference fixes 0.4 (ForwardProjection.cpp:444)
    int spectrumStep = 1;
    int specEnergyNum = 0;
    ScintilltorInfo mScintilltorInfo;
};

struct GridInfo {
    int interspcaceWidth = 0; // um
    int gridStripsWidth = 0;  // um
    float h = 0;              // mm
    float FD = 0;             // mm
    std::string gridStripMaterial;
    float gridStripDensity = 0;
    std::string gridStripMAPath;
    std::string interspaceMaterial;
    float interspaceDensity = 0;
    std::string interspaceMAPath;
    float gridStripLinearAtten = 0;  // 1/mm
    float interspaceLinearAtten = 0; // 1/mm
};

struct CTScanSystemInfo {
    float totalTime = 0.0f;
    int rotatedDirection = 1;
    float thetaStep = 0, dHalfLU = 0, dHalfLV = 0, FOVR = 0, FOVH = 0;
    float pHalfX = 0, pHalfY = 0, pHalfZ = 0;
    float pSizeX = 0, pSizeY = 0, pSizeZ = 0;
    float dx = 0, dy = 0, dz = 0;
    size_t intNum = 0;
    float spectrumVal = 0;
    float phantomMAtten = 0;
};

// Sample coordinates of the reference's device layer (Source/DataType.h:120-126:
// one FP32 position per sample, K per ray).  The projector sets every ray up
// from its geometry instead and never fills these; the struct is kept so that
// code written against the reference compiles.
struct Coordinate {
    float *imgIntX = nullptr, *imgIntY = nullptr, *imgIntZ = nullptr;
};

// Host arrays of one scan (the reference's h_mForwardProj bag).
struct PolyForwardProj {
    uint8 *phantomStructure = nullptr;
    float *phantom = nullptr;
    float *phantomMassAtten = nullptr;
    float *spectrumNormal = nullptr;
    float *grid = nullptr;
    float *gridLinearAtten = nullptr;
    float *interspaceLinearAtten = nullptr;
    float *I = nullptr, *I0 = nullptr, *IAbsorb = nullptr;
    float *proj = nullptr;
    float *scintillatorLineAtten = nullptr;
    float *scintillatorPerThickness = nullptr;
    float *detResponse = nullptr;
    float *foSpOffsetU = nullptr, *foSpOffsetV = nullptr;
};
typedef PolyForwardProj ForwardProj;

struct FilePath {
    std::string spectrumPath;
    std::string gridPath;
    std::string phantomPath;
    std::string phantomMAttenPath;
    std::string outputFolder;
    std::string scintillatorPath;
    std::string convKernelGridPath;
};

struct PhantomMaterial {
    int materialNum = 0;
    std::string structureMarkPath;
    float *density = nullptr;
    std::string *massAttenPath = nullptr;
};

enum ScanMode { TestI0, GridNoResponse, GridResponse, NoGridNoResponse, NoGridResponse };

namespace ctfp {
// Failure of a host step (missing file, bad configuration, device error).
// The reference printed and called exit(0); these classes throw instead.
struct Error : std::exception {
    int code;
    std::string msg;
    Error(int c, std::string m) : code(c), msg(std::move(m)) {}
    const char *what() const noexcept override { return msg.c_str(); }
};
// Reads the first n whitespace-separated values of a one-column table
// (InputData/Readme.txt:1-2); backslashes in the path are accepted.
void readTable(const std::string &path, int n, float *out, const char *what);
std::string nativePath(const std::string &p);
} // namespace ctfp
