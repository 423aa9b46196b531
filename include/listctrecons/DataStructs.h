// DataStructs.h -- the data model of Include/DataStructs.h:9-334 with the same
// classes, constructors and getters, free of CUDA: geometry converts to the
// C-ABI structs of lct.h, device buffers are GPUMemoryCtrl over lct_device_*.
//
// Deviations (each keeps the reference's observable results):
//  * sizes are 64-bit (the reference's int products overflow for large scans,
//    Src/DataReader.cpp:14, Src/ImageWriter.cpp:20);
//  * BlockInfo's block index is per object instead of a static member
//    (Include/DataStructs.h:278), so two ReconData objects do not interfere;
//  * CTDataGPU also holds the block's fan-beam views on the device (the
//    reference keeps them in a CUDA array bound to a texture reference).
#pragma once
#include <cstddef>
#include <string>

#include "lct.h"

class FileParams;
template <class T> class GPUMemoryCtrl;

class CTGeom
{
private:
    float orbit, orbitStart;
    int ns, nt;
    float ws, wt;
    int nAngle, nAngleAdd;
    float ds, dt;
    float pitch;
    float sourceZ0;
    float dso;
    int nProjTurn;
    int zDir;
    int addProjections;

public:
    CTGeom() {}
    CTGeom(const float orbit, const float orbitStart, const int ns, const int nt, const int na,
           const int naa, const float ds, const float dt, const float ws, const float wt,
           const float pitch, const float sourceZ, const float dso, const int nProjTurn,
           const int zDir, const int addProjections)
        : orbit(orbit), orbitStart(orbitStart), ns(ns), nt(nt), ws(ws), wt(wt), nAngle(na),
          nAngleAdd(naa), ds(ds), dt(dt), pitch(pitch), sourceZ0(sourceZ), dso(dso),
          nProjTurn(nProjTurn), zDir(zDir), addProjections(addProjections)
    {
    }
    explicit CTGeom(const lct_ct_geom &g)
        : CTGeom(g.orbit, g.orbit_start, g.ns, g.nt, g.n_angle, g.n_angle_add, g.ds, g.dt, g.ws,
                 g.wt, g.pitch, g.source_z0, g.dso, g.n_proj_turn, g.z_dir, g.add_projections)
    {
    }
    lct_ct_geom to_c() const
    {
        lct_ct_geom g;
        g.orbit = orbit;
        g.orbit_start = orbitStart;
        g.ns = ns;
        g.nt = nt;
        g.ws = ws;
        g.wt = wt;
        g.n_angle = nAngle;
        g.n_angle_add = nAngleAdd;
        g.ds = ds;
        g.dt = dt;
        g.pitch = pitch;
        g.source_z0 = sourceZ0;
        g.dso = dso;
        g.n_proj_turn = nProjTurn;
        g.z_dir = zDir;
        g.add_projections = addProjections;
        return g;
    }
    float orbit_() const { return this->orbit; }
    float orbit_start() const { return this->orbitStart; }
    int ns_() const { return this->ns; }
    int nt_() const { return this->nt; }
    float ws_() const { return this->ws; }
    float wt_() const { return this->wt; }
    int n_angle() const { return this->nAngle; }
    int n_angle_add() const { return this->nAngleAdd; }
    float ds_() const { return this->ds; }
    float dt_() const { return this->dt; }
    float pitch_() const { return this->pitch; }
    float source_z0() const { return this->sourceZ0; }
    float dso_() const { return this->dso; }
    int n_proj_turn() const { return this->nProjTurn; }
    int z_dir() const { return this->zDir; }
    int add_projections() const { return this->addProjections; }
};

class ImgGeom
{
private:
    int nx, ny, nz;
    float wx, wy, wz;
    float dx, dy, dz;

public:
    ImgGeom() {}
    ImgGeom(int nx, int ny, int nz, float dx, float dy, float dz, float offX, float offY,
            float offZ)
        : nx(nx), ny(ny), nz(nz), wx((nx - 1) / 2.0f + offX), wy((ny - 1) / 2.0f + offY),
          wz(offZ), dx(dx), dy(dy), dz(dz)
    {
    }
    lct_img_geom to_c() const
    {
        lct_img_geom g;
        g.nx = nx;
        g.ny = ny;
        g.nz = nz;
        g.wx = wx;
        g.wy = wy;
        g.wz = wz;
        g.dx = dx;
        g.dy = dy;
        g.dz = dz;
        return g;
    }
    int nx_() const { return this->nx; }
    int ny_() const { return this->ny; }
    int nz_() const { return this->nz; }
    float wx_() const { return this->wx; }
    float wy_() const { return this->wy; }
    float wz_() const { return this->wz; }
    float dx_() const { return this->dx; }
    float dy_() const { return this->dy; }
    float dz_() const { return this->dz; }
};

class FileInfo
{
private:
    std::string inFileDir;
    std::string dataFileName;
    std::string outputDir;
    std::string outputName;

public:
    FileInfo(const std::string inFileDir, const std::string dataFileName,
             const std::string outputDir, const std::string outputName)
        : inFileDir(inFileDir), dataFileName(dataFileName), outputDir(outputDir),
          outputName(outputName)
    {
    }
    std::string in_file_dir() const { return this->inFileDir; }
    std::string data_file_name() const { return this->dataFileName; }
    std::string output_dir() const { return this->outputDir; }
    std::string output_name() const { return this->outputName; }
};

class CTDataCPU
{
private:
    float *proj;
    float *rebin;
    float *image;
    float *tubeAngles;
    float *tablePositions;
    CTDataCPU(const CTDataCPU &) = delete;
    CTDataCPU &operator=(const CTDataCPU &) = delete;

public:
    CTDataCPU() : proj(nullptr), rebin(nullptr), image(nullptr), tubeAngles(nullptr), tablePositions(nullptr) {}
    CTDataCPU(const float *tubeAngle, const float *tablePosition, const size_t sizeProj,
              const size_t sizeRebin, const size_t sizeImage, const int nReadings);
    ~CTDataCPU();
    float *proj_() const { return this->proj; }
    float *rebin_() const { return this->rebin; }
    float *image_() const { return this->image; }
    float *tube_angles() const { return this->tubeAngles; }
    float *table_positions() const { return this->tablePositions; }
};

class CTDataGPU
{
private:
    GPUMemoryCtrl<float> *fanGpu;
    GPUMemoryCtrl<float> *projGpu;
    GPUMemoryCtrl<float> *imgGpu;
    GPUMemoryCtrl<float> *betasGpu;
    CTDataGPU(const CTDataGPU &) = delete;
    CTDataGPU &operator=(const CTDataGPU &) = delete;

public:
    CTDataGPU();
    ~CTDataGPU();
    GPUMemoryCtrl<float> *fan_gpu() const { return this->fanGpu; }
    GPUMemoryCtrl<float> *proj_gpu() const { return this->projGpu; }
    GPUMemoryCtrl<float> *img_gpu() const { return this->imgGpu; }
    GPUMemoryCtrl<float> *betas_gpu() const { return this->betasGpu; }
};

class ReconInfo
{
private:
    int nReadings;
    float dataBeginPos;
    float dataEndPos;
    float allowBegin;
    float allowEnd;
    float startPos;
    float endPos;
    float reconStartPos;
    float reconEndPos;
    int nSlicesRecon;
    int nSlicesBlock;
    int nBlocks;
    int dataOffset;

public:
    ReconInfo(const int nReadings, const float dataBeginPos, const float dataEndPos,
              const float allowBegin, const float allowEnd, const float startPos,
              const float endPos, const float reconStartPos, const float reconEndPos,
              const int nSlicesRecon, const int nSlicesBlock, const int nBlocks,
              const int dataOffset)
        : nReadings(nReadings), dataBeginPos(dataBeginPos), dataEndPos(dataEndPos),
          allowBegin(allowBegin), allowEnd(allowEnd), startPos(startPos), endPos(endPos),
          reconStartPos(reconStartPos), reconEndPos(reconEndPos), nSlicesRecon(nSlicesRecon),
          nSlicesBlock(nSlicesBlock), nBlocks(nBlocks), dataOffset(dataOffset)
    {
    }
    int n_readings() const { return this->nReadings; }
    float data_begin_pos() const { return this->dataBeginPos; }
    float data_end_pos() const { return this->dataEndPos; }
    float allow_begin() const { return this->allowBegin; }
    float allow_end() const { return this->allowEnd; }
    float start_pos() const { return this->startPos; }
    float end_pos() const { return this->endPos; }
    float recon_start_pos() const { return this->reconStartPos; }
    float recon_end_pos() const { return this->reconEndPos; }
    int n_slices_recon() const { return this->nSlicesRecon; }
    int n_slices_block() const { return this->nSlicesBlock; }
    int n_blocks() const { return this->nBlocks; }
    int data_offset() const { return this->dataOffset; }
};

class BlockInfo
{
private:
    int blockId;
    int blockProjStart;
    int blockProjEnd;
    float blockStart;
    float blockEnd;
    int blockSliceStart;
    int blockSliceEnd;
    const lct_plan *plan;

public:
    BlockInfo(const int blockProjStart, const int blockProjEnd, const float blockStart,
              const float blockEnd, const int blockSliceStart, const int blockSliceEnd,
              const lct_plan *plan)
        : blockId(0), blockProjStart(blockProjStart), blockProjEnd(blockProjEnd),
          blockStart(blockStart), blockEnd(blockEnd), blockSliceStart(blockSliceStart),
          blockSliceEnd(blockSliceEnd), plan(plan)
    {
    }
    int block_idx() const { return blockId; }
    int block_proj_start() const { return this->blockProjStart; }
    int block_proj_end() const { return this->blockProjEnd; }
    float block_start() const { return this->blockStart; }
    float block_end() const { return this->blockEnd; }
    int block_slice_start() const { return this->blockSliceStart; }
    int block_slice_end() const { return this->blockSliceEnd; }
    void update(const ReconInfo *ri, const CTGeom *ctg, const ImgGeom *ig,
                const CTDataCPU *dataCPU);
};

class ReconData
{
private:
    CTGeom *ctg;
    ImgGeom *ig;
    FileInfo *fi;
    CTDataCPU *dataCPU;
    CTDataGPU *dataGPU;
    ReconInfo *ri;
    BlockInfo *bi;
    lct_plan plan;
    void convert_from_file_params(const FileParams &fp);
    ReconData(const ReconData &) = delete;
    ReconData &operator=(const ReconData &) = delete;

public:
    ReconData(const FileParams &fp) { convert_from_file_params(fp); }
    ~ReconData();
    CTGeom *ctg_() const { return this->ctg; }
    ImgGeom *ig_() const { return this->ig; }
    FileInfo *fi_() const { return this->fi; }
    CTDataCPU *data_cpu() const { return this->dataCPU; }
    CTDataGPU *data_gpu() const { return this->dataGPU; }
    ReconInfo *ri_() const { return this->ri; }
    BlockInfo *bi_() const { return this->bi; }
    // The C-ABI plan (lct_plan) the block windows are derived from.
    const lct_plan *plan_() const { return &this->plan; }
};
