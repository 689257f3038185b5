// block.h -- Block and Neighbor of the HiLeMMS API (reference Src/block.h:
// 55-104, ranges of Src/block.cpp:14-140, including the "OP_3D" spelling
// that keeps TopBack/TopFront on the full y range).
#ifndef BLOCK_H
#define BLOCK_H
#include <map>
#include <string>
#include <vector>

#include "boundary_host_device.h"
#include "flowfield_host_device.h"
#include "ops_seq_v2.h"
#include "type.h"

struct Neighbor {
    int blockId{0};
    BoundarySurface surface;
    VertexType type;
};

class Block {
   public:
    Block(const int blockId, const std::string& blockName,
          const std::vector<int>& blockSize);
    int ID() const { return id; }
    const std::string& Name() const { return name; }
    const ops_block& Get() const { return block; }
    ops_block& Get() { return block; }
    const std::vector<int>& Size() const { return size; }
    const int* pSize() const { return size.data(); }
    const std::vector<int>& WholeRange() const { return wholeRange; }
    const std::vector<int>& BulkRange() const { return bulkRange; }
    const std::map<BoundarySurface, std::vector<int>>& BoundarySurfaceRange() const {
        return boundarySurfaceRange;
    }
    const std::map<BoundarySurface, Neighbor>& Neighbors() const { return neighbors; }
    void AddNeighbor(BoundarySurface surface, const Neighbor& neighbor);

   private:
    ops_block block{nullptr};
#ifdef OPS_3D
    const int spaceDim{3};
#else
    const int spaceDim{2};
#endif
    std::string name;
    int id{0};
    std::vector<int> size;
    std::map<BoundarySurface, std::vector<int>> boundarySurfaceRange;
    std::vector<int> wholeRange;
    std::vector<int> bulkRange;
    std::map<BoundarySurface, Neighbor> neighbors;
    int RangeStart(const int axis, const BoundarySurface surface) const;
    int RangeEnd(const int axis, const BoundarySurface surface) const;
};
using BlockGroup = std::map<int, Block>;
#endif
