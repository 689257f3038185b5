// field.h -- Field<T>: one host mirror (ops_dat) per block, as the
// reference's Src/field.h:51-99 surface (at/[] give the block's ops_dat).
// Device storage of the populations and macros lives in libmplbhip; these
// mirrors carry what Apps read and write on the host.
#ifndef FIELD_H
#define FIELD_H
#include <map>
#include <string>

#include "block.h"
#include "ops_seq_v2.h"
#include "type.h"

template <typename T>
class Field {
   public:
    Field(const std::string& varName, const int dataDim = 1, const int halo = 1)
        : name(varName), dim(dataDim), haloDepth(halo) {}
    Field(const char* varName, const int dataDim = 1, const int halo = 1)
        : name(varName), dim(dataDim), haloDepth(halo) {}
    void CreateFieldFromScratch(const BlockGroup& blocks) {
        for (const auto& idBlock : blocks) CreateFieldFromScratch(idBlock.second);
    }
    void CreateFieldFromScratch(const Block& block) {
        std::string dname{name + "_" + block.Name()};
        int sz[3] = {1, 1, 1};
        for (size_t a = 0; a < block.Size().size(); a++) sz[a] = block.Size()[a];
        data[block.ID()] = ops_decl_dat_host(block.Get(), block.ID(), dim, sz, haloDepth,
                                             sizeof(T), sizeof(T) == 8 ? "double" : "int",
                                             dname.c_str());
    }
    void SetDataDim(const int dataDim) { dim = dataDim; }
    void SetDataHalo(const int halo) { haloDepth = halo; }
    int HaloDepth() const { return haloDepth; }
    int DataDim() const { return dim; }
    const std::string& Name() const { return name; }
    ops_dat& at(int blockIdx) { return data.at(blockIdx); }
    const ops_dat& at(int blockIdx) const { return data.at(blockIdx); }
    ops_dat& operator[](int blockIdx) { return at(blockIdx); }
    const ops_dat& operator[](int blockIdx) const { return at(blockIdx); }
    bool Has(int blockIdx) const { return data.count(blockIdx) != 0; }
    // host value at node (i,j,k) of block b, component d
    T& Value(int b, int i, int j, int k = 0, int d = 0) {
        ops_dat dat = data.at(b);
        return reinterpret_cast<T*>(dat->data.data())[dat->offset(i, j, k) + d];
    }

   private:
    std::map<int, ops_dat> data;
    std::string name;
    int dim{1};
    int haloDepth{1};
};

using RealField = Field<Real>;
using IntField = Field<int>;
using IntFieldGroup = std::map<int, IntField>;
using RealFieldGroup = std::map<int, RealField>;
#endif
