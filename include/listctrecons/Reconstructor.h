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
//
// With slab_blocks > 0 the range is streamed instead: slabs of slab_blocks
// 32-slice blocks, each reading only the views it needs
// (lct_slab_views / lct_reconstruct_slabs).  With checkpointing on, every
// finished slab is written into the output image file at its place and
// "<output>.ckpt" records how many blocks are done; a later run with the same
// plan resumes after them (the file is removed when all blocks are done).
// stop_after_slabs > 0 ends the process (exit status 3) after that many
// slabs of this run, as an interrupted run would.
class FDKResidentReconstructor : public Reconstructor
{
private:
    lct_recon_opts opts;
    int slab_blocks = 0;
    bool checkpoint = false;
    int stop_after_slabs = 0;

public:
    explicit FDKResidentReconstructor(const lct_recon_opts &o) : opts(o) {}
    void set_streaming(int blocks_per_slab, bool with_checkpoint, int stop_after = 0)
    {
        slab_blocks = blocks_per_slab;
        checkpoint = with_checkpoint;
        stop_after_slabs = stop_after;
    }
    void reconstruct(const ReconData *mr);
    const lct_recon_opts &options() const { return opts; }
    int slabs_per_step() const { return slab_blocks; }
    bool checkpointing() const { return checkpoint; }
};
