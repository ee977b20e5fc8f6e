#include "Grid.h"

#include <cstdio>
#include <filesystem>
#include <fstream>

#include "ctfp.h"

namespace {
ctfp_grid_params params(const GridAndDetectorSystem &m)
{
    ctfp_grid_params g;
    g.d_num = (int64_t)m.dNums;
    g.d_size_um = m.dSize;
    g.interspace_width = m.interspcaceWidth;
    g.strip_width = m.gridStripsWidth;
    g.height_mm = m.h;
    g.focal_mm = m.FD;
    return g;
}
void check(int rc)
{
    if (rc != CTFP_OK) throw ctfp::Error(rc, ctfp_last_error());
}
} // namespace

Grid::Grid(GridAndDetectorSystem s) : m(s) {}
Grid::Grid(float *gridOutput, GridAndDetectorSystem s) : m(s), grid(gridOutput) {}
Grid::~Grid() {}

void Grid::defGrid()
{
    if (!grid) {
        own.assign(m.dNums, 1.0f);
        grid = own.data();
    }
    // a transparent evaluation validates the layout once
    float zero = 0.0f;
    std::vector<float> t(m.dNums);
    ctfp_grid_params g = params(m);
    check(ctfp_grid_transmission(&g, 1, &zero, &zero, t.data()));
}

void Grid::updataGrid(float gridStripLinearAtten, float interspaceLinearAtten)
{
    if (!grid) defGrid();
    m.gridStripLinearAtten = gridStripLinearAtten;
    m.interspaceLinearAtten = interspaceLinearAtten;
    ctfp_grid_params g = params(m);
    check(ctfp_grid_transmission(&g, 1, &gridStripLinearAtten, &interspaceLinearAtten, grid));
}

void Grid::transmission(int n_energy, const float *mu_strip, const float *mu_inter, float *T) const
{
    ctfp_grid_params g = params(m);
    check(ctfp_grid_transmission(&g, n_energy, mu_strip, mu_inter, T));
}

void Grid::getPeriodIncex(std::vector<size_t> &gridPeriodIndex)
{
    if (!grid) defGrid();
    std::vector<int64_t> idx(m.dNums + 2);
    int64_t n = 0;
    check(ctfp_grid_period_index((int64_t)m.dNums, grid, idx.data(), (int64_t)idx.size(), &n));
    periodIndex.assign(idx.begin(), idx.begin() + n);
    gridPeriodIndex = periodIndex;
}

void Grid::savePeriodIndex(const std::string &path) const
{
    std::ofstream f(path);
    if (!f.is_open()) throw ctfp::Error(CTFP_ERR_IO, "cannot write " + path);
    for (size_t i : periodIndex) f << i << "\n";
}

namespace {
void writeFloats(const std::string &dir, const std::string &name, const float *p, size_t n)
{
    std::filesystem::create_directories(dir);
    const std::string path = dir + "/" + name;
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw ctfp::Error(CTFP_ERR_IO, "cannot write " + path);
    f.write((const char *)p, (std::streamsize)(n * sizeof(float)));
    if (!f) throw ctfp::Error(CTFP_ERR_IO, "short write to " + path);
}
} // namespace

void Grid::showGridH()
{
    ctfp_grid_params g = params(m);
    int64_t n = 0;
    check(ctfp_grid_unit_heights(&g, nullptr, 0, &n));
    std::vector<float> H((size_t)n);
    check(ctfp_grid_unit_heights(&g, H.data(), n, &n));
    char name[512];
    std::snprintf(name, sizeof name,
                  "Grid_dNums%zux%zu_dSize%dum_S%s_IS%s_gridD%dum_gridd%dum_h%.2fmm_float_FD%.1fmm_Hight%lld.raw",
                  m.dNums, m.dNums, (int)m.dSize, m.gridStripMaterial.c_str(), m.interspaceMaterial.c_str(),
                  m.interspcaceWidth, m.gridStripsWidth, (double)m.h, (double)m.FD, (long long)n);
    writeFloats("GridImageData", name, H.data(), H.size());
}

void Grid::showPeriodIncex()
{
    std::vector<size_t> idx;
    getPeriodIncex(idx);
    savePeriodIndex("GridPeriodIndex_" + std::to_string(m.dNums) + ".txt");
}

void Grid::creatBackground()
{
    if (!grid) defGrid();
    std::vector<float> I0(m.dNums * m.dNums);
    for (size_t v = 0; v < m.dNums; v++)
        for (size_t u = 0; u < m.dNums; u++) I0[v * m.dNums + u] = grid[u] * (float)m.mI0Photons;
    char name[512];
    std::snprintf(name, sizeof name, "Background(Grid)_dNums%zux%zu_dSize%dum_gridD%dum_gridd%dum_h%.2f_Photon%zu.raw",
                  m.dNums, m.dNums, (int)m.dSize, m.interspcaceWidth, m.gridStripsWidth, (double)(m.h * 1000),
                  m.mI0Photons);
    writeFloats("AddIntensityImageData", name, I0.data(), I0.size());
}
