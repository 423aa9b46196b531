// Driver for the reference's own .prm parser (Src/FileParams.cpp, compiled
// from /root/reference by ref.mk): prints every parsed value, one per line,
// in the format of `LISTCTRecons --prm`.  Test infrastructure only.
#include <cstdio>

#include "FileParams.h"

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    FileParams fp(argv[1]);
    std::printf("RawDataDir %s\nDataInterpFile %s\nViewAngleFile %s\nCouchPosFile %s\n",
                fp.raw_file_dir().c_str(), fp.data_interp_file().c_str(),
                fp.view_angle_file().c_str(), fp.couch_pos_file().c_str());
    std::printf("OutputDir %s\nOutputName %s\n", fp.output_dir().c_str(), fp.output_name().c_str());
    std::printf("Ns %d\nNt %d\nDs %.9g\nDt %.9g\nPitchValue %.9g\nNx %d\nNy %d\nDx %.9g\nDz %.9g\n",
                fp.ns(), fp.nt(), fp.ds(), fp.dt(), fp.pitch_value(), fp.nx(), fp.ny(), fp.dx(),
                fp.dz());
    std::printf("OffsetX %.9g\nOffsetY %.9g\nReadings %d\nViewOffset %d\nDataOffset %d\n",
                fp.offset_x(), fp.offset_y(), fp.readings(), fp.view_offset(), fp.data_offset());
    std::printf("RSrcToIso %.9g\nCentralChannel %.9g\nNProjTurn %d\n", fp.r_src_to_iso(),
                fp.central_channel(), fp.n_proj_turn());
    return 0;
}
