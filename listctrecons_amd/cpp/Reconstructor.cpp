// Reconstructor.cpp -- FDKReconstructor (Src/Reconstructor.cpp:10-49), the
// HBM-resident reconstructor, and ReconstrctorAbstractFactory
// (Src/ReconstructorAbstractFactory.cpp:11-33).
#include "Reconstructor.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "Backprojector.h"
#include "DataReader.h"
#include "DataStructs.h"
#include "Filter.h"
#include "GPUMemoryCtrl.h"
#include "PreWeighting.h"
#include "Rebiner.h"
#include "ReconstructorAbstractFactory.h"

FDKReconstructor::FDKReconstructor(Rebiner *rbn, Filter *flt, Backprojector *bp, PreWeighting *pw)
    : rbn(rbn), flt(flt), bp(bp), dr(new RawDataReader()), pw(pw)
{
}

FDKReconstructor::~FDKReconstructor()
{
    delete rbn;
    delete flt;
    delete bp;
    delete dr;
    delete pw;
}

void FDKReconstructor::reconstruct(const ReconData *mr)
{
    std::cout << "Start reconstruction..." << std::endl;
    for (int i = 0; i < mr->ri_()->n_blocks(); i++) {
        std::cout << std::endl << "Start processing block " << i << "..." << std::endl;
        std::cout << "Reading data..." << std::endl;
        dr->readData(mr);
        std::cout << "Rebining..." << std::endl;
        rbn->rebin(mr);
        std::cout << "Weighting..." << std::endl;
        pw->weighting(mr);
        std::cout << "Filtering..." << std::endl;
        flt->filter(mr);
        std::cout << "Backprojecting..." << std::endl;
        bp->backproject(mr);
        mr->bi_()->update(mr->ri_(), mr->ctg_(), mr->ig_(), mr->data_cpu());
        std::cout << "Block " << i << " finished." << std::endl;
    }
}

void FDKResidentReconstructor::reconstruct(const ReconData *mr)
{
    std::cout << "Start reconstruction (HBM-resident views)..." << std::endl;
    lct_plan plan = *mr->plan_();
    const float *couch = mr->data_cpu()->table_positions();
    const float *angles = mr->data_cpu()->tube_angles();
    // file views [lo, hi) cover every block's window
    long lo = 1L << 40, hi = 0;
    for (int b = 0; b < plan.n_blocks; b++) {
        int s = 0;
        lct_block_info(&plan, couch, b, &s, nullptr, nullptr, nullptr, nullptr);
        lo = std::min(lo, long(s) + plan.data_offset);
        hi = std::max(hi, long(s) + plan.data_offset + plan.ctg.n_angle_add);
    }
    if (lo < 0) {
        std::cout << "[ERROR] projection window starts before the data file" << std::endl;
        std::exit(1);
    }
    const size_t view = size_t(plan.ctg.ns) * plan.ctg.nt;
    const size_t n = view * size_t(hi - lo);
    std::vector<float> host(n);
    const std::string path = mr->fi_()->in_file_dir() + mr->fi_()->data_file_name();
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f || std::fseek(f, long(lo * long(view) * long(sizeof(float))), SEEK_SET) != 0 ||
        std::fread(host.data(), sizeof(float), n, f) != n) {
        std::cout << "[ERROR] could not read " << n << " values from " << path << std::endl;
        std::exit(1);
    }
    std::fclose(f);
    GPUMemoryCtrl<float> fan, vol;
    const size_t n_vol = size_t(plan.ig.nx) * plan.ig.ny * plan.n_slices_recon;
    if (!fan.allocate_memory(n) || !fan.copy_from_CPU(host.data(), n) || !vol.allocate_memory(n_vol)) {
        std::cout << "[ERROR] " << lct_last_error() << std::endl;
        std::exit(1);
    }
    plan.data_offset -= int(lo);  // the device buffer starts at file view `lo`
    if (lct_reconstruct(lct_default_ctx(), &plan, &opts, angles, couch, fan.get_pointer(),
                        hi - lo, vol.get_pointer()) != LCT_OK ||
        lct_synchronize(lct_default_ctx()) != LCT_OK) {
        std::cout << "[ERROR] " << lct_last_error() << std::endl;
        std::exit(1);
    }
    vol.copy_to_CPU(mr->data_cpu()->image_(), n_vol);
    std::cout << "Reconstruction finished (" << plan.n_blocks << " blocks)." << std::endl;
}

namespace {
bool has(const std::string &s, const char *t) { return s.find(t) != std::string::npos; }
}

// The reference tests `str.find("fdk")`, which is true for every string that
// does not START with "fdk" (Src/ReconstructorAbstractFactory.cpp:13); that
// behaviour is kept.  Its default string "Helicalfdkramprebingpu" builds the
// reference stage set.  Further words in the string select the variants:
//   "piori"        PiOriRebiner + PiOriPreWeighting + PiOriBackprojectorGPU
//   "null"         NullRebinerGPU
//   "hann"/"shepp" apodisation window of RampFilterGPU
//   "resident"     FDKResidentReconstructor, FP32 storage
//   "f16"/"bf16"   FDKResidentReconstructor, f16 / bf16 filtered projections
//                  with the BF16x3 Toeplitz MFMA ramp filter
Reconstructor *ReconstrctorAbstractFactory::createReconstructor(const std::string &str)
{
    if (str.find("fdk")) {
        std::string s = str;
        std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return std::tolower(c); });
        SmoothWindow win = SmoothWindow::NONE;
        if (has(s, "hann")) win = SmoothWindow::HANN;
        else if (has(s, "shepp")) win = SmoothWindow::SHEPP_LOGAN;
        const bool piori = has(s, "piori");
        const bool null_rebin = has(s, "null");
        if (has(s, "resident") || has(s, "f16")) {
            lct_recon_opts o{};
            o.pipeline = LCT_PIPE_RESIDENT;
            o.rebin_kind = piori ? LCT_REBIN_PI_ORI : (null_rebin ? LCT_REBIN_NULL : LCT_REBIN_FAN2PAR);
            o.weight_kind = piori ? LCT_WEIGHT_PI_ORI : LCT_WEIGHT_NONE;
            o.bp_kind = piori ? LCT_BP_PI_ORI : LCT_BP_HELICAL;
            o.window = int(win);
            o.storage = has(s, "bf16") ? LCT_BF16 : (has(s, "f16") ? LCT_F16 : LCT_F32);
            o.filter_algo = o.storage == LCT_F32 ? LCT_FILTER_DIRECT_F32 : LCT_FILTER_TOEPLITZ_BF16X3;
            o.interp = LCT_INTERP_LDS;
            return new FDKResidentReconstructor(o);
        }
        Rebiner *rbn = nullptr;
        PreWeighting *pw = nullptr;
        Backprojector *bp = nullptr;
        Filter *flt = new RampFilterGPU(win);
        if (piori) {
            rbn = new PiOriRebiner();
            pw = new PiOriPreWeighting();
            bp = new PiOriBackprojectorGPU();
        } else {
            rbn = null_rebin ? static_cast<Rebiner *>(new NullRebinerGPU())
                             : static_cast<Rebiner *>(new HelicalConeBeamFan2ParRebinerGPU());
            pw = new NULLPreWeighting();
            bp = new HelicalConeBeamFanRebinBackprojectorGPU();
        }
        return new FDKReconstructor(rbn, flt, bp, pw);
    }
    std::cout << "Error happens when building reconstructor." << std::endl;
    std::exit(1);
    return nullptr;
}
