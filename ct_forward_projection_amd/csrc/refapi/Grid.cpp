#include "Grid.h"

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
