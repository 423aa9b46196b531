// Filter.h -- Include/Filter.h:6-24 (template method: filter() calls the
// private virtual calculate_filter()).  RampFilterGPU takes an optional
// apodisation window; the default is the reference's plain Ram-Lak.
#pragma once

#include "SmoothWindow.h"

class ReconData;
class CTGeom;

class Filter
{
private:
    virtual void calculate_filter(float *h, const CTGeom *ctg) = 0;

public:
    Filter() {}
    virtual ~Filter() {}
    void filter(const ReconData *mr);
};

class RampFilterGPU : public Filter
{
private:
    SmoothWindow window;
    void calculate_filter(float *h, const CTGeom *ctg);

public:
    RampFilterGPU(SmoothWindow w = SmoothWindow::NONE) : window(w) {}
    ~RampFilterGPU() {}
};
