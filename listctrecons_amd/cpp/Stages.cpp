// Stages.cpp -- the reference's stage plug-ins over the C ABI:
//   Rebiner      CUDA/Rebiner.cu:31-80, :105-154, :181-233
//   PreWeighting CUDA/PreWeighting.cu:31-44
//   Filter       CUDA/Filter.cu:52-73 + Src/Filter.cpp:8-20
//   Backprojector CUDA/Backprojector.cu:101-169, :252-307
// Each call processes the current block held by ReconData and synchronises
// before returning, as the reference does.  A device error prints the
// library's message and exits (the reference printed and carried on).
#include <cstdlib>
#include <iostream>
#include <vector>

#include "Backprojector.h"
#include "DataStructs.h"
#include "Filter.h"
#include "GPUMemoryCtrl.h"
#include "PreWeighting.h"
#include "Rebiner.h"

namespace {

void ok(int rc, const char *what)
{
    if (rc != LCT_OK) {
        std::cout << "[ERROR] " << what << ": " << lct_last_error() << std::endl;
        std::exit(1);
    }
}

void need(bool b, const char *what)
{
    if (!b) {
        std::cout << "[ERROR] " << what << std::endl;
        std::exit(1);
    }
}

// Uploads the block's fan-beam views and rebins them into proj_gpu.
void rebin_block(const ReconData *mr, int kind)
{
    const CTGeom *ctg = mr->ctg_();
    CTDataGPU *gpu = mr->data_gpu();
    const size_t view = size_t(ctg->ns_()) * ctg->nt_();
    const size_t n_fan = view * ctg->n_angle_add();
    if (gpu->fan_gpu()->size() == 0) need(gpu->fan_gpu()->allocate_memory(n_fan), "allocate fan views");
    need(gpu->fan_gpu()->copy_from_CPU(mr->data_cpu()->proj_(), n_fan), "copy fan views to the device");
    if (gpu->proj_gpu()->size() == 0)
        need(gpu->proj_gpu()->allocate_memory(view * ctg->n_angle()), "allocate proj_gpu");
    const lct_ct_geom g = ctg->to_c();
    ok(lct_rebin(lct_default_ctx(), kind, LCT_WEIGHT_NONE, &g, gpu->fan_gpu()->get_pointer(),
                 gpu->proj_gpu()->get_pointer()),
       "rebin");
    ok(lct_synchronize(lct_default_ctx()), "rebin");
    std::cout << "Rebin finished" << std::endl;
}

void backproject_block(const ReconData *mr, int kind)
{
    const CTGeom *ctg = mr->ctg_();
    const ImgGeom *ig = mr->ig_();
    CTDataGPU *gpu = mr->data_gpu();
    const size_t n_img = size_t(ig->nx_()) * ig->ny_() * ig->nz_();
    if (gpu->img_gpu()->size() == 0) need(gpu->img_gpu()->allocate_memory(n_img), "allocate img_gpu");
    // betas = tubeAngles[blockProjStart + add ...] (Backprojector.cu:145), with
    // the conjugate-index tail of clipped windows (npt/2 + 1 more views).
    const int nb = ctg->n_angle() + ctg->n_proj_turn() / 2 + 1;
    std::vector<float> betas(nb, 0.0f);
    const long first = long(mr->bi_()->block_proj_start()) + ctg->add_projections();
    for (int i = 0; i < nb; i++)
        if (first + i >= 0 && first + i < mr->ri_()->n_readings())
            betas[i] = mr->data_cpu()->tube_angles()[first + i];
    const lct_ct_geom g = ctg->to_c();
    const lct_img_geom gi = ig->to_c();
    ok(lct_backproject(lct_default_ctx(), kind, &g, &gi, gpu->proj_gpu()->get_pointer(),
                       betas.data(), gpu->img_gpu()->get_pointer()),
       "backproject");
    ok(lct_synchronize(lct_default_ctx()), "backproject");
    std::cout << "backprojection finished." << std::endl;
    need(gpu->img_gpu()->copy_to_CPU(mr->data_cpu()->image_() + n_img * mr->bi_()->block_idx(), n_img),
         "copy the block image to the host");
    // the reference frees both per block (Backprojector.cu:138, :167)
    gpu->proj_gpu()->free_memory();
    gpu->img_gpu()->free_memory();
}

}  // namespace

void NullRebinerGPU::rebin(const ReconData *mr) { rebin_block(mr, LCT_REBIN_NULL); }

void HelicalConeBeamFan2ParRebinerGPU::rebin(const ReconData *mr) { rebin_block(mr, LCT_REBIN_FAN2PAR); }

void PiOriRebiner::rebin(const ReconData *mr) { rebin_block(mr, LCT_REBIN_PI_ORI); }

int PiOriPreWeighting::weighting(const ReconData *mr)
{
    const lct_ct_geom g = mr->ctg_()->to_c();
    if (lct_weighting(lct_default_ctx(), LCT_WEIGHT_PI_ORI, &g,
                      mr->data_gpu()->proj_gpu()->get_pointer()) != LCT_OK)
        return -1;
    return lct_synchronize(lct_default_ctx()) == LCT_OK ? 1 : -1;
}

void Filter::filter(const ReconData *mr)
{
    const CTGeom *ctg = mr->ctg_();
    CTDataGPU *gpu = mr->data_gpu();
    std::vector<float> h(2 * size_t(ctg->ns_()));
    calculate_filter(h.data(), ctg);
    const lct_ct_geom g = ctg->to_c();
    const long rows = long(ctg->n_angle()) * ctg->nt_();
    ok(lct_filter(lct_default_ctx(), &g, h.data(), gpu->proj_gpu()->get_pointer(), rows), "filter");
    ok(lct_synchronize(lct_default_ctx()), "filter");
    // the filtered block goes to the host too (Filter.cu:69); rebin.raw dumps it
    need(gpu->proj_gpu()->copy_to_CPU(mr->data_cpu()->rebin_(), size_t(rows) * ctg->ns_()),
         "copy the filtered block to the host");
    std::cout << "Filter finished" << std::endl;
}

void RampFilterGPU::calculate_filter(float *h, const CTGeom *ctg)
{
    ok(lct_ramp_taps(int(window), ctg->ns_(), ctg->ds_(), h), "ramp taps");
}

void HelicalConeBeamFanRebinBackprojectorGPU::backproject(const ReconData *mr)
{
    backproject_block(mr, LCT_BP_HELICAL);
}

void PiOriBackprojectorGPU::backproject(const ReconData *mr) { backproject_block(mr, LCT_BP_PI_ORI); }
