// DataIO.cpp -- RawDataReader (Src/DataReader.cpp:8-19) and RawImageWriter
// (Src/ImageWriter.cpp:9-43) with 64-bit offsets and sizes, and checked I/O.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "DataReader.h"
#include "DataStructs.h"
#include "ImageWriter.h"

void RawDataReader::readData(const ReconData *mr)
{
    const std::string filename = mr->fi_()->in_file_dir() + mr->fi_()->data_file_name();
    FILE *fp = std::fopen(filename.c_str(), "rb");
    if (!fp) {
        std::perror("Raw data file not found.");
        std::exit(1);
    }
    const CTGeom *g = mr->ctg_();
    const long long view = (long long)g->ns_() * g->nt_();
    const long long first = (long long)mr->bi_()->block_proj_start() + mr->ri_()->data_offset();
    const size_t count = size_t(view) * size_t(g->n_angle_add());
    if (first < 0 || fseeko(fp, off_t(first * view * (long long)sizeof(float)), SEEK_SET) != 0 ||
        std::fread(mr->data_cpu()->proj_(), sizeof(float), count, fp) != count) {
        std::cout << "Projection file " << filename << " does not hold views " << first << " to "
                  << first + g->n_angle_add() << "." << std::endl;
        std::fclose(fp);
        std::exit(1);
    }
    std::fclose(fp);
}

void RawImageWriter::writeImage(const ReconData *mr)
{
    const std::string filename = mr->fi_()->output_dir() + mr->fi_()->output_name();
    std::ofstream file(filename.c_str(), std::ios::out | std::ios::binary);
    if (!file.is_open()) {
        std::cout << "Opening image file " << mr->fi_()->output_dir() << mr->fi_()->output_name()
                  << " failed." << std::endl;
        std::exit(1);
    }
    std::cout << "Writing image to " << mr->fi_()->output_dir() << mr->fi_()->output_name() << std::endl;
    const size_t image_bytes =
        size_t(mr->ig_()->nx_()) * mr->ig_()->ny_() * mr->ri_()->n_slices_recon() * sizeof(float);
    file.write(reinterpret_cast<const char *>(mr->data_cpu()->image_()), std::streamsize(image_bytes));
    file.close();

    const std::string filename2 = mr->fi_()->output_dir() + "rebin.raw";
    std::ofstream file2(filename2.c_str(), std::ios::out | std::ios::binary);
    if (!file2.is_open()) {
        std::cout << "Opening image file " << mr->fi_()->output_dir() << "rebin.raw failed." << std::endl;
        std::exit(1);
    }
    std::cout << "Writing image to " << mr->fi_()->output_dir() << "rebin.raw" << std::endl;
    const size_t rebin_bytes =
        size_t(mr->ctg_()->ns_()) * mr->ctg_()->nt_() * mr->ctg_()->n_angle() * sizeof(float);
    file2.write(reinterpret_cast<const char *>(mr->data_cpu()->rebin_()), std::streamsize(rebin_bytes));
    file2.close();
    std::cout << "Writing finished." << std::endl;
}
