// hdf5_io.h -- checkpoint files of the HiLeMMS layer (see hdf5_io.cpp for
// the layout, which follows the reference's OPS files).
#ifndef HILEMMS_HDF5_IO_H
#define HILEMMS_HDF5_IO_H
#include <string>

#include "ops_seq_v2.h"

namespace hilemms_h5 {
// <case>_<block>_T<step>.h5 (Src/field.h:180-205)
std::string file_name(const std::string& caseName, const std::string& blockName,
                      long long timeStep);
// write (or replace) one block's field with its halo into the file
void write_dat(const std::string& file, const std::string& blockName, int blockIndex,
               int spaceDim, const ops_dat d);
// read one block's field back into its host mirror (shapes must match)
void read_dat(const std::string& file, const std::string& blockName, int spaceDim, ops_dat d);
}  // namespace hilemms_h5
#endif
