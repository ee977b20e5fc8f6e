// Grid.h -- focused anti-scatter grid with the interface of the reference's
// Grid class (Source/Grid.h:42-158).  The transmission of every pixel is the
// mean of its 1-um lamella units, computed by ctfp_grid_transmission.
#pragma once
#include <cstddef>
#include <string>
#include <vector>

#include "DataType.h"

struct GridAndDetectorSystem {
    size_t dNums = 0;           // pixels along u
    float dSize = 0;            // um
    int interspcaceWidth = 0;   // um
    int gridStripsWidth = 0;    // um
    float h = 0;                // mm
    float FD = 0;               // mm
    std::string gridStripMaterial;
    float gridStripDensity = 0; // g/cm^3
    std::string gridStripMAPath;
    std::string interspaceMaterial;
    float interspaceDensity = 0;
    std::string interspaceMAPath;
    float gridStripLinearAtten = 0;  // 1/mm
    float interspaceLinearAtten = 0; // 1/mm
    size_t mI0Photons = 0;
    size_t angleNums = 0;
};

class Grid {
  public:
    Grid(GridAndDetectorSystem mGridAndDetectorSystem);
    Grid(float *gridOutput, GridAndDetectorSystem mGridAndDetectorSystem);
    ~Grid();

    // Validates the layout (Grid.cpp:373-405).
    void defGrid();
    // Transmission of one energy into the bound output [dNums]
    // (Grid.cpp:408-443); attenuations in 1/mm.
    void updataGrid(float gridStripLinearAtten, float interspaceLinearAtten);
    // All energies at once: T[e][dNums].
    void transmission(int n_energy, const float *mu_strip, const float *mu_inter, float *T) const;
    // 0, every strict local minimum of the current row, dNums-1 (Grid.cpp:531-536, 605-620).
    void getPeriodIncex(std::vector<size_t> &gridPeriodIndex);
    // Writes the period indices, one per line, to `path`.
    void savePeriodIndex(const std::string &path) const;
    // The slant heights (um) of the 1-um lamella units, as FP32 raw, to
    // GridImageData/Grid_dNums..._Hight<units>.raw under the working
    // directory (Grid.cpp:506-523).
    void showGridH();
    // Retrieves the period indices of the current row and writes them to
    // GridPeriodIndex_<dNums>.txt in the working directory (Grid.cpp:525-529).
    void showPeriodIncex();
    // Flat field behind the grid for a square dNums x dNums detector:
    // row[u] * mI0Photons in every row, as FP32 raw, to
    // AddIntensityImageData/Background(Grid)_... (Grid.cpp:553-561, 583-592).
    void creatBackground();

  private:
    GridAndDetectorSystem m;
    float *grid = nullptr;        // bound output of updataGrid
    std::vector<float> own;       // used when no output is bound
    std::vector<size_t> periodIndex;
};
