// main.cpp -- the LISTCTRecons CLI of Src/main.cpp:10-49.
//   LISTCTRecons <PRM filename> [reconstructor string]
//   LISTCTRecons --plan <PRM filename>     print the planner's result (no GPU)
// The reconstructor string defaults to the prm's optional ReconType key, then
// to the reference's "Helicalfdkramprebingpu".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "DataStructs.h"
#include "FileParams.h"
#include "ImageWriter.h"
#include "ReconstructorAbstractFactory.h"
#include "Reconstructor.h"

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::cout << "Usage LISTCTRecons <PRM filename>" << std::endl;
        return 1;
    }
    if (std::strcmp(argv[1], "--prm") == 0 && argc > 2) {
        FileParams fp(argv[2]);
        std::printf("RawDataDir %s\nDataInterpFile %s\nViewAngleFile %s\nCouchPosFile %s\n",
                    fp.raw_file_dir().c_str(), fp.data_interp_file().c_str(),
                    fp.view_angle_file().c_str(), fp.couch_pos_file().c_str());
        std::printf("OutputDir %s\nOutputName %s\n", fp.output_dir().c_str(), fp.output_name().c_str());
        std::printf("Ns %d\nNt %d\nDs %.9g\nDt %.9g\nPitchValue %.9g\nNx %d\nNy %d\nDx %.9g\nDz %.9g\n",
                    fp.ns(), fp.nt(), fp.ds(), fp.dt(), fp.pitch_value(), fp.nx(), fp.ny(),
                    fp.dx(), fp.dz());
        std::printf("OffsetX %.9g\nOffsetY %.9g\nReadings %d\nViewOffset %d\nDataOffset %d\n",
                    fp.offset_x(), fp.offset_y(), fp.readings(), fp.view_offset(), fp.data_offset());
        std::printf("RSrcToIso %.9g\nCentralChannel %.9g\nNProjTurn %d\n", fp.r_src_to_iso(),
                    fp.central_channel(), fp.n_proj_turn());
        return 0;
    }
    if (std::strcmp(argv[1], "--plan") == 0) {
        if (argc < 3) {
            std::cout << "Usage LISTCTRecons --plan <PRM filename>" << std::endl;
            return 1;
        }
        FileParams fp(argv[2]);
        ReconData mr(fp);
        const lct_plan *p = mr.plan_();
        std::printf("add %d n_angle_add %d n_angle %d n_slices_recon %d n_blocks %d z_dir %d orbit %.9g\n",
                    p->ctg.add_projections, p->ctg.n_angle_add, p->ctg.n_angle, p->n_slices_recon,
                    p->n_blocks, p->ctg.z_dir, p->ctg.orbit);
        for (int b = 0; b < p->n_blocks; b++) {
            int s = 0, e = 0;
            lct_block_info(p, mr.data_cpu()->table_positions(), b, &s, &e, nullptr, nullptr, nullptr);
            std::printf("block %d proj_start %d proj_end %d\n", b, s, e);
        }
        return 0;
    }
    const std::string prmFile = argv[1];
    // optional: [reconstructor string] [--stop-after-slabs N]
    std::string arg_type;
    int stop_after = 0;
    for (int i = 2; i < argc; i++) {
        if (std::strcmp(argv[i], "--stop-after-slabs") == 0 && i + 1 < argc)
            stop_after = std::atoi(argv[++i]);
        else
            arg_type = argv[i];
    }
    std::cout << "Reading prm file..." << std::endl;
    FileParams *fp = new FileParams(prmFile);
    std::cout << "Converting parameters..." << std::endl;
    ReconData *mr = new ReconData(*fp);
    std::string recon_type = "Helicalfdkramprebingpu";
    if (!fp->recon_type().empty()) recon_type = fp->recon_type();
    if (!arg_type.empty()) recon_type = arg_type;
    std::cout << "Creating reconstructor..." << std::endl;
    ReconstrctorAbstractFactory *rFac = new ReconstrctorAbstractFactory();
    Reconstructor *recon = rFac->createReconstructor(recon_type);
    if (stop_after > 0) {
        auto *res = dynamic_cast<FDKResidentReconstructor *>(recon);
        if (!res) {
            std::cout << "--stop-after-slabs needs a streamed (slab / ckpt) reconstructor" << std::endl;
            return 1;
        }
        res->set_streaming(res->slabs_per_step(), res->checkpointing(), stop_after);
    }
    recon->reconstruct(mr);
    ImageWriter *imWriter = new RawImageWriter();
    imWriter->writeImage(mr);
    delete fp;
    delete mr;
    delete rFac;
    delete recon;
    delete imWriter;
    return 0;
}
