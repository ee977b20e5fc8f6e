// PhantomObject.h -- label volume + per-material tables, with the interface of
// the reference's PhantomObject (Source/PhantomObject.h:9-47).
#pragma once
#include <vector>

#include "DataType.h"

class PhantomObject {
  public:
    PhantomObject(const CTScanParas mCTScanParas, const PhantomMaterial mPhantomMaterial);
    ~PhantomObject();

    // Reads the uint8 label volume and every material's mass-attenuation table.
    void readData();
    // Multi-material linear attenuation (1/mm) of energy bin specIndex,
    // PhantomObject.cpp:38-58 (labels above materialNum keep the buffer value).
    void updataLineAtten(float *lineAttenCoef, size_t specIndex);
    // Single material: (mu/rho)/10 of bin specIndex, PhantomObject.cpp:60-63.
    void updataMassAtten(float &massAttenCoef, size_t specIndex);
    // Single material density = label * rho0, PhantomObject.cpp:65-71.
    void computeDesity(float *density);

    // Accessors for the device path.
    const uint8 *labels() const { return _structureMark.data(); }
    int materialNum() const { return _mPhantomMaterial.materialNum; }
    const float *densities() const { return _density.data(); }
    float massAtten(size_t specIndex, int material) const
    {
        return _massAttenCoef[specIndex * _mPhantomMaterial.materialNum + material];
    }

  private:
    void _readPhantomStructure();
    void _readMassAtten();

    PhantomMaterial _mPhantomMaterial;
    std::vector<float> _density;
    std::vector<std::string> _massAttenPath;
    std::vector<uint8> _structureMark;
    std::vector<float> _massAttenCoef; // [energy][material], cm^2/g
    int _specEnergyNum;
    size_t _pNumX, _pNumY, _pNumZ;
};
