// Reconstructor.h -- Include/Reconstructor.h:10-31, plus the HBM-resident
// reconstructor that runs the whole range through lct_reconstruct.
#pragma once

#include "lct.h"

class Rebiner;
class Filter;
class Backprojector;
class DataReader;
class ReconData;
class PreWeighting;

class Reconstructor
{
public:
    Reconstructor() {}
    virtual ~Reconstructor() {}
    virtual void reconstruct(const ReconData *mr) = 0;
};

class FDKReconstructor : public Reconstructor
{
private:
    Rebiner *rbn;
    Filter *flt;
    Backprojector *bp;
    DataReader *dr;
    PreWeighting *pw;

public:
    FDKReconstructor(Rebiner *rbn, Filter *flt, Backprojector *bp, PreWeighting *pw);
    ~FDKReconstructor();
    void reconstruct(const ReconData *mr);
};

// Whole planned range in one lct_reconstruct call: the projection range is
// read once into HBM, every view rebinned and filtered once, and all blocks
// backprojected from the resident filtered views.
class FDKResidentReconstructor : public Reconstructor
{
private:
    lct_recon_opts opts;

public:
    explicit FDKResidentReconstructor(const lct_recon_opts &o) : opts(o) {}
    void reconstruct(const ReconData *mr);
    const lct_recon_opts &options() const { return opts; }
};
