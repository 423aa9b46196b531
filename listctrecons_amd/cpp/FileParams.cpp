// FileParams.cpp -- the .prm parser of Src/FileParams.cpp:10-142: whitespace
// tokens "Key: value", all 23 keys required ("TableFeed:" is an alias of
// "PitchValue:"), unknown tokens ignored, exit(1) on a missing file or key.
// The optional "ReconType:" key (ignored by the reference) selects the
// factory string for the CLI.
#include "FileParams.h"

#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>

namespace {
const char *kTokenNames[23] = {"RawDataDir", "DataInterpFile", "ViewAngleFile", "CouchPosFile",
                               "OutputDir", "OutputName", "Ns", "Nt", "Ds", "Dt", "PitchValue",
                               "Nx", "Ny", "Dx", "Dz", "OffsetX", "OffsetY", "Readings",
                               "ViewOffset", "DataOffset", "RSrcToIso", "CentralChannel",
                               "NProjTurn"};
}

bool FileParams::enough_data(const bool haveData[])
{
    for (int i = 0; i < 23; i++) {
        if (!haveData[i]) {
            std::cout << "Parameter not found in prm file:" << kTokenNames[i] << std::endl;
            std::exit(1);
        }
    }
    return true;
}

void FileParams::readPrmFile(const std::string prmFileName)
{
    bool have[23] = {false};
    std::ifstream file(prmFileName.c_str());
    if (!file.is_open()) {
        std::cout << "Parameter file not found. " << std::endl;
        std::exit(1);
    }
    std::string token;
    while (file >> token) {
        if (token == "RawDataDir:") { file >> RawFileDir; have[0] = true; }
        else if (token == "DataInterpFile:") { file >> DataInterpFile; have[1] = true; }
        else if (token == "ViewAngleFile:") { file >> ViewAngleFile; have[2] = true; }
        else if (token == "CouchPosFile:") { file >> CouchPosFile; have[3] = true; }
        else if (token == "OutputDir:") { file >> OutputDir; have[4] = true; }
        else if (token == "OutputName:") { file >> OutputName; have[5] = true; }
        else if (token == "Ns:") { file >> Ns; have[6] = true; }
        else if (token == "Nt:") { file >> Nt; have[7] = true; }
        else if (token == "Ds:") { file >> Ds; have[8] = true; }
        else if (token == "Dt:") { file >> Dt; have[9] = true; }
        else if (token == "PitchValue:" || token == "TableFeed:") { file >> PitchValue; have[10] = true; }
        else if (token == "Nx:") { file >> Nx; have[11] = true; }
        else if (token == "Ny:") { file >> Ny; have[12] = true; }
        else if (token == "Dx:") { file >> Dx; have[13] = true; }
        else if (token == "Dz:") { file >> Dz; have[14] = true; }
        else if (token == "OffsetX:") { file >> OffsetX; have[15] = true; }
        else if (token == "OffsetY:") { file >> OffsetY; have[16] = true; }
        else if (token == "Readings:") { file >> Readings; have[17] = true; }
        else if (token == "ViewOffset:") { file >> ViewOffset; have[18] = true; }
        else if (token == "DataOffset:") { file >> DataOffset; have[19] = true; }
        else if (token == "RSrcToIso:") { file >> RSrcToIso; have[20] = true; }
        else if (token == "CentralChannel:") { file >> CentralChannel; have[21] = true; }
        else if (token == "NProjTurn:") { file >> NProjTurn; have[22] = true; }
        else if (token == "ReconType:") { file >> ReconType; }
    }
    enough_data(have);
}

FileParams::FileParams(const std::string prmFileName) { this->readPrmFile(prmFileName); }
