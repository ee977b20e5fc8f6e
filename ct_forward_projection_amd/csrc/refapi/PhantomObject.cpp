#include "PhantomObject.h"

#include <fstream>
#include <iostream>

#include "ctfp.h"

PhantomObject::PhantomObject(const CTScanParas p, const PhantomMaterial m)
    : _mPhantomMaterial(m), _specEnergyNum(p.specEnergyNum), _pNumX(p.pNumX), _pNumY(p.pNumY),
      _pNumZ(p.pNumZ)
{
    for (int i = 0; i < m.materialNum; i++) {
        _density.push_back(m.density ? m.density[i] : 0.0f);
        _massAttenPath.push_back(m.massAttenPath ? m.massAttenPath[i] : std::string());
    }
}

PhantomObject::~PhantomObject() {}

void PhantomObject::readData()
{
    _readPhantomStructure();
    _readMassAtten();
}

void PhantomObject::_readPhantomStructure()
{
    size_t n = _pNumX * _pNumY * _pNumZ;
    _structureMark.assign(n, 0);
    std::ifstream f(ctfp::nativePath(_mPhantomMaterial.structureMarkPath), std::ios::binary);
    if (!f.is_open())
        throw ctfp::Error(CTFP_ERR_IO, "cannot open the phantom structure file " +
                                           _mPhantomMaterial.structureMarkPath);
    f.read((char *)_structureMark.data(), (std::streamsize)n);
    if ((size_t)f.gcount() != n)
        throw ctfp::Error(CTFP_ERR_IO, "phantom structure file " + _mPhantomMaterial.structureMarkPath +
                                           " holds fewer than " + std::to_string(n) + " voxels");
}

void PhantomObject::_readMassAtten()
{
    int M = _mPhantomMaterial.materialNum;
    _massAttenCoef.assign((size_t)_specEnergyNum * M, 0.0f);
    std::vector<float> col(_specEnergyNum);
    for (int m = 0; m < M; m++) {
        ctfp::readTable(_massAttenPath[m], _specEnergyNum, col.data(),
                        ("mass attenuation of material " + std::to_string(m + 1)).c_str());
        for (int e = 0; e < _specEnergyNum; e++) _massAttenCoef[(size_t)e * M + m] = col[e];
    }
}

void PhantomObject::updataLineAtten(float *lineAttenCoef, size_t specIndex)
{
    const int M = _mPhantomMaterial.materialNum;
    for (size_t i = 0; i < _structureMark.size(); i++) {
        unsigned lab = _structureMark[i];
        if (lab == 0)
            lineAttenCoef[i] = 0;
        else if (lab <= (unsigned)M)
            lineAttenCoef[i] = _massAttenCoef[specIndex * M + lab - 1] * _density[lab - 1] / 10.f;
    }
}

void PhantomObject::updataMassAtten(float &massAttenCoef, size_t specIndex)
{
    massAttenCoef = _massAttenCoef[specIndex * _mPhantomMaterial.materialNum] / 10.f;
}

void PhantomObject::computeDesity(float *density)
{
    for (size_t i = 0; i < _structureMark.size(); i++) density[i] = _structureMark[i] * _density[0];
}
