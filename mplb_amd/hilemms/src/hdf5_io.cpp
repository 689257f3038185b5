// hdf5_io.cpp -- checkpoint files of the HiLeMMS layer in the layout the
// reference writes through OPS (Field::WriteToHDF5 and CreateFieldFromFile,
// Src/field.h:166-205) and reads back with Src/PostProcess.py:107-196:
//
//   file     <case>_<block name>_T<time step>.h5, one per block and step
//   group    /<block name>           attributes ops_type = "ops_block",
//                                     dims, index
//   dataset  /<block name>/<field>_<block name>
//            the field with its halo, i fastest and the field's components
//            fastest of all: shape (nz + 2h, ny + 2h, (nx + 2h) dim) in 3D,
//            (ny + 2h, (nx + 2h) dim) in 2D; attributes ops_type = "ops_dat",
//            block, block_index, dim, size, d_m, d_p, base, type
//
// Every field of a block at one step goes into the same file (appended, a
// dataset of the same name replaced).
#include "hdf5_io.h"

#include <hdf5.h>
#include <hdf5_hl.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace hilemms_h5 {

namespace {

[[noreturn]] void h5die(const char* what, const std::string& file) {
    std::fprintf(stderr, "Error! HDF5: %s (%s)\n", what, file.c_str());
    std::fflush(stderr);
    std::abort();
}

hid_t open_or_create(const std::string& file) {
    H5E_auto2_t func;
    void* data;
    H5Eget_auto2(H5E_DEFAULT, &func, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    hid_t f = H5Fopen(file.c_str(), H5F_ACC_RDWR, H5P_DEFAULT);
    if (f < 0) f = H5Fcreate(file.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    H5Eset_auto2(H5E_DEFAULT, func, data);
    if (f < 0) h5die("cannot open or create the file", file);
    return f;
}

// dataset shape of a dat: space axes reversed, components folded into x
std::vector<hsize_t> shape_of(const ops_dat d, int space_dim) {
    std::vector<hsize_t> shp;
    for (int a = space_dim - 1; a >= 0; a--) {
        hsize_t n = (hsize_t)(d->size[a] - d->d_m[a] + d->d_p[a]);
        if (a == 0) n *= (hsize_t)d->dim;
        shp.push_back(n);
    }
    return shp;
}

hid_t mem_type(const ops_dat d) {
    return d->elem_size == 8 ? H5T_NATIVE_DOUBLE : H5T_NATIVE_INT;
}

}  // namespace

std::string file_name(const std::string& caseName, const std::string& blockName,
                      long long timeStep) {
    return caseName + "_" + blockName + "_T" + std::to_string(timeStep) + ".h5";
}

void write_dat(const std::string& file, const std::string& blockName, int blockIndex,
               int spaceDim, const ops_dat d) {
    const hid_t f = open_or_create(file);
    // the block group (ops_fetch_block_hdf5_file)
    hid_t g;
    if (H5Lexists(f, blockName.c_str(), H5P_DEFAULT) > 0) {
        g = H5Gopen2(f, blockName.c_str(), H5P_DEFAULT);
    } else {
        g = H5Gcreate2(f, blockName.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (g < 0) h5die("cannot create the block group", file);
        H5LTset_attribute_string(f, blockName.c_str(), "ops_type", "ops_block");
        H5LTset_attribute_int(f, blockName.c_str(), "dims", &spaceDim, 1);
        H5LTset_attribute_int(f, blockName.c_str(), "index", &blockIndex, 1);
    }
    if (g < 0) h5die("cannot open the block group", file);
    // the dataset (ops_fetch_dat_hdf5_file)
    if (H5Lexists(g, d->name.c_str(), H5P_DEFAULT) > 0) H5Ldelete(g, d->name.c_str(), H5P_DEFAULT);
    const std::vector<hsize_t> shp = shape_of(d, spaceDim);
    const hid_t sp = H5Screate_simple((int)shp.size(), shp.data(), nullptr);
    const hid_t ds = H5Dcreate2(g, d->name.c_str(), mem_type(d), sp, H5P_DEFAULT, H5P_DEFAULT,
                                H5P_DEFAULT);
    if (ds < 0) h5die("cannot create the dataset", file);
    if (H5Dwrite(ds, mem_type(d), H5S_ALL, H5S_ALL, H5P_DEFAULT, d->data.data()) < 0)
        h5die("cannot write the dataset", file);
    H5Dclose(ds);
    H5Sclose(sp);
    const std::string path = blockName + "/" + d->name;
    int size[3], dm[3], dp[3], base[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) {
        size[a] = d->size[a] - d->d_m[a] + d->d_p[a];  // OPS counts the halo in
        dm[a] = d->d_m[a];                             // a dat's size
        dp[a] = d->d_p[a];
    }
    H5LTset_attribute_string(f, path.c_str(), "ops_type", "ops_dat");
    H5LTset_attribute_string(f, path.c_str(), "block", blockName.c_str());
    H5LTset_attribute_int(f, path.c_str(), "block_index", &blockIndex, 1);
    H5LTset_attribute_int(f, path.c_str(), "dim", &d->dim, 1);
    H5LTset_attribute_int(f, path.c_str(), "size", size, (size_t)spaceDim);
    H5LTset_attribute_int(f, path.c_str(), "d_m", dm, (size_t)spaceDim);
    H5LTset_attribute_int(f, path.c_str(), "d_p", dp, (size_t)spaceDim);
    H5LTset_attribute_int(f, path.c_str(), "base", base, (size_t)spaceDim);
    H5LTset_attribute_string(f, path.c_str(), "type", d->elem_size == 8 ? "double" : "int");
    H5Gclose(g);
    H5Fclose(f);
}

void read_dat(const std::string& file, const std::string& blockName, int spaceDim, ops_dat d) {
    const hid_t f = H5Fopen(file.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) h5die("cannot open the restart file", file);
    const std::string path = blockName + "/" + d->name;
    const hid_t ds = H5Dopen2(f, path.c_str(), H5P_DEFAULT);
    if (ds < 0) h5die(("no dataset " + path).c_str(), file);
    const hid_t sp = H5Dget_space(ds);
    const int rank = H5Sget_simple_extent_ndims(sp);
    std::vector<hsize_t> shp((size_t)std::max(rank, 0));
    H5Sget_simple_extent_dims(sp, shp.data(), nullptr);
    if (shp != shape_of(d, spaceDim)) h5die(("dataset " + path + " has another shape").c_str(), file);
    if (H5Dread(ds, mem_type(d), H5S_ALL, H5S_ALL, H5P_DEFAULT, d->data.data()) < 0)
        h5die(("cannot read " + path).c_str(), file);
    H5Sclose(sp);
    H5Dclose(ds);
    H5Fclose(f);
    d->dirty = true;
}

}  // namespace hilemms_h5
