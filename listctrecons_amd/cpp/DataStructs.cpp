// DataStructs.cpp -- ReconData set-up (Src/DataStructs.cpp:110-250) and
// BlockInfo::update (:77-108) over the C-ABI planner.
#include "DataStructs.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "Common.h"
#include "FileParams.h"
#include "GPUMemoryCtrl.h"

namespace {
lct_ctx *g_ctx = nullptr;

void destroy_ctx()
{
    if (g_ctx) lct_ctx_destroy(g_ctx);
    g_ctx = nullptr;
}

// Reads `count` float32 values after skipping `skip` values; exit(1) as the
// reference does when the file is missing (Src/DataStructs.cpp:139-156).
std::vector<float> read_floats(const std::string &path, int skip, int count)
{
    std::ifstream f(path.c_str(), std::ios::in | std::ios::binary);
    if (!f.is_open()) {
        std::perror("Raw data file not found.");
        std::exit(1);
    }
    std::vector<float> all(size_t(skip) + size_t(count));
    f.read(reinterpret_cast<char *>(all.data()), std::streamsize(all.size() * sizeof(float)));
    if (size_t(f.gcount()) != all.size() * sizeof(float)) {
        std::cout << "File " << path << " holds fewer than " << all.size() << " float32 values."
                  << std::endl;
        std::exit(1);
    }
    return std::vector<float>(all.begin() + skip, all.end());
}
}  // namespace

lct_ctx *lct_default_ctx()
{
    if (!g_ctx) {
        const char *dev = std::getenv("LCT_DEVICE");
        if (lct_ctx_create(dev ? std::atoi(dev) : 0, &g_ctx) != LCT_OK) {
            std::cout << "[ERROR] " << lct_last_error() << std::endl;
            std::exit(1);
        }
        std::atexit(destroy_ctx);
    }
    return g_ctx;
}

CTDataCPU::CTDataCPU(const float *tubeAngle, const float *tablePosition, const size_t sizeProj,
                     const size_t sizeRebin, const size_t sizeImage, const int nReadings)
{
    proj = new float[sizeProj]();
    rebin = new float[sizeRebin]();
    image = new float[sizeImage]();
    tubeAngles = new float[nReadings]();
    tablePositions = new float[nReadings]();
    std::memcpy(tubeAngles, tubeAngle, size_t(nReadings) * sizeof(float));
    std::memcpy(tablePositions, tablePosition, size_t(nReadings) * sizeof(float));
}

CTDataCPU::~CTDataCPU()
{
    delete[] proj;
    delete[] rebin;
    delete[] image;
    delete[] tubeAngles;
    delete[] tablePositions;
}

CTDataGPU::CTDataGPU()
    : fanGpu(new GPUMemoryCtrl<float>()), projGpu(new GPUMemoryCtrl<float>()),
      imgGpu(new GPUMemoryCtrl<float>()), betasGpu(new GPUMemoryCtrl<float>())
{
}

CTDataGPU::~CTDataGPU()
{
    delete fanGpu;
    delete projGpu;
    delete imgGpu;
    delete betasGpu;
}

void BlockInfo::update(const ReconInfo *, const CTGeom *, const ImgGeom *, const CTDataCPU *dataCPU)
{
    blockId++;
    int s = 0, e = 0;
    float bs = 0.0f, be = 0.0f;
    if (lct_block_info(plan, dataCPU->table_positions(), blockId, &s, &e, &bs, &be, nullptr) !=
        LCT_OK) {
        std::cout << "[ERROR] " << lct_last_error() << std::endl;
        std::exit(1);
    }
    this->blockStart = bs;
    this->blockEnd = be;
    this->blockSliceStart = BLOCK_SLICES * blockId;
    this->blockSliceEnd = blockSliceStart + BLOCK_SLICES - 1;
    this->blockProjStart = s;
    this->blockProjEnd = e;
}

ReconData::~ReconData()
{
    delete bi;
    delete ig;
    delete ctg;
    delete dataCPU;
    delete dataGPU;
    delete fi;
    delete ri;
}

void ReconData::convert_from_file_params(const FileParams &fp)
{
    const std::string datapath = fp.raw_file_dir() + fp.data_interp_file();
    {
        std::ifstream raw(datapath.c_str(), std::ios::in | std::ios::binary);
        if (!raw.is_open()) {
            std::perror("Raw data file not found.");
            std::exit(1);
        }
    }
    std::vector<float> angles = read_floats(fp.raw_file_dir() + fp.view_angle_file(),
                                            fp.view_offset(), fp.readings());
    std::vector<float> couch = read_floats(fp.raw_file_dir() + fp.couch_pos_file(),
                                           fp.view_offset(), fp.readings());

    lct_prm p;
    p.ns = fp.ns();
    p.nt = fp.nt();
    p.ds = fp.ds();
    p.dt = fp.dt();
    p.pitch_value = fp.pitch_value();
    p.nx = fp.nx();
    p.ny = fp.ny();
    p.dx = fp.dx();
    p.dz = fp.dz();
    p.offset_x = fp.offset_x();
    p.offset_y = fp.offset_y();
    p.readings = fp.readings();
    p.view_offset = fp.view_offset();
    p.data_offset = fp.data_offset();
    p.r_src_to_iso = fp.r_src_to_iso();
    p.central_channel = fp.central_channel();
    p.n_proj_turn = fp.n_proj_turn();
    if (lct_plan_helical(&p, angles.data(), couch.data(), &plan) != LCT_OK) {
        std::printf("%s\n", lct_last_error());
        std::exit(1);
    }
    std::printf("Reconstructing couch position from %f to %f, %d slices in total.\n",
                plan.recon_start_pos, plan.recon_end_pos, plan.n_slices_recon);

    ctg = new CTGeom(plan.ctg);
    const lct_img_geom &g = plan.ig;
    ig = new ImgGeom(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, fp.offset_x(), fp.offset_y(), 0.0f);
    fi = new FileInfo(fp.raw_file_dir(), fp.data_interp_file(), fp.output_dir(), fp.output_name());
    ri = new ReconInfo(plan.n_readings, plan.data_begin_pos, plan.data_end_pos, plan.allow_begin,
                       plan.allow_end, plan.start_pos, plan.end_pos, plan.recon_start_pos,
                       plan.recon_end_pos, plan.n_slices_recon, plan.n_slices_block,
                       plan.n_blocks, plan.data_offset);
    const size_t view = size_t(ctg->ns_()) * ctg->nt_();
    dataCPU = new CTDataCPU(angles.data(), couch.data(), view * ctg->n_angle_add(),
                            view * ctg->n_angle(),
                            size_t(ig->nx_()) * ig->ny_() * ri->n_slices_recon(), fp.readings());
    dataGPU = new CTDataGPU();
    int s = 0, e = 0;
    float bs = 0.0f, be = 0.0f;
    lct_block_info(&plan, couch.data(), 0, &s, &e, &bs, &be, nullptr);
    bi = new BlockInfo(s, e, bs, be, 0, plan.n_slices_block - 1, &plan);
}
