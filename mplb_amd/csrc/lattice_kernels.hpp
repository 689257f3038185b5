// lattice_kernels.hpp -- parameter blocks shared by the host runtime and the
// gfx950 kernels.
#pragma once
#include <cstdint>

#include "plan.hpp"

namespace mplbhip {

constexpr int kTX = 64;   // one wavefront along x
constexpr int kTY = 4;    // four wavefronts per workgroup along y
constexpr int kNAcc = 9;  // residual: 4 squared differences, 4 squares, mass
// FP16 storage holds h = g / s_q with s_q = w_q / kF16Scale: g / w_q is
// O(1e-1) for |u| ~ 0.05, so h stays in FP16's normal range with ~11
// significant bits
constexpr double kF16Scale = 1024.0;
// doubles in the file-scope scratch array of the allocation-free entry
// point (8 MiB per GPU): kNAcc partials per workgroup of a whole-lattice
// launch plus kNAcc totals per block and for the case
constexpr long long kScratchDoubles = 1LL << 20;

// One block as laid out in HBM: a one node halo on each axis (no z halo in
// 2D), rows padded so that x = 0 is 64-byte aligned; the Q populations of a
// z plane follow each other (pitch sq = one padded plane), and z planes
// follow at pitch sz = Q * sq.  Element (q,x,y,z): base + z*sz + q*sq + y*sx + x.
struct BlockParams {
    int32_t n[3];
    int32_t sx;          // row pitch (elements)
    int32_t sz;          // z-slab pitch (all populations of one plane)
    int32_t sq;          // population pitch within a slab
    int64_t base;        // element offset of node (0,0,0), population 0
    int32_t tiles_x, tiles_y;
    int32_t tile_start;  // first logical workgroup of this block
    int32_t prog_base;   // first region program of this block
    int64_t mbase;       // offset of this block in the macro arrays
    int64_t fbase;       // offset of this block in host-layout f arrays
    int64_t persist_base;
    double g[3];         // acceleration (code units)
    double dt;           // time step of this block
    double tau, omega, one_minus_omega, tau_omega;
    float fomega, fone_minus_omega, fforce, fhalf_dtg[3];
    int32_t active;      // 0 freezes the block (batched instances)
    // split step: the bulk kernel covers the interior region (z planes 1 ..
    // n-2 in 3D) in tiles_x * tiles_y tiles per plane; the boundary kernel
    // walks the block's boundary segments, 256 nodes per workgroup
    int32_t bulk_tile_start;  // first bulk workgroup of this block
    // fold level 3: the bulk grid spans every z plane (3D), so the z-face
    // planes' simple regions run in it too; its first workgroup
    int32_t bulk_tile_start3;
    // boundary workgroups and segments per fold level: [0] every boundary
    // region, [1] without the simple x-face interiors, [2] without every
    // simple region of the planes z = 1 .. n-2, [3] without every simple
    // region (those run inside the bulk kernel)
    int32_t bnd_wg_start[4];
    int32_t seg_start[4], seg_count[4];
    int32_t xgroup;           // tiles_x when this block's bulk tiles can be
                              // dealt a row per XCD, else 0
    int32_t xgroup3;          // the same for the fold-3 bulk grid
    int32_t tiles_x2;         // FP16 two-node bulk kernel: 128-node tiles per row
    int32_t h2_tile_start;    // its first workgroup of this block
    int32_t rowwg;            // 1: tiles_x is a multiple of kTY and n[1] of kTY,
                              // so a bulk workgroup can take kTY tiles of one row
    int32_t has_obst;         // the block has obstacle link masks
    // divisors of the bulk tile index, each as a multiply-high-and-shift
    // pair (fast_div): [0] tiles_x, [1] tiles_y, [2] n[1], [3] tiles_x / kTY
    uint32_t div_m[4];
    int32_t div_sh[4];
    // the bulk kernel's population accesses as raw buffer accesses: one
    // resource per array, the lane's node offset, and a scalar byte offset
    // per population.  Loads use the resource base prev - pull_shift
    // elements, which keeps every offset non-negative: pull_off[q] =
    // (q sq - delta_q + pull_shift) * elem_bytes with delta_q = c_qx +
    // c_qy sx + c_qz sz; stores use next itself, put_off[q] = q sq * elem_bytes.
    int32_t pull_shift;
    uint32_t pull_off[19];
    uint32_t put_off[19];
};

// n / d for 0 <= n < 2^31 as floor(n * m / 2^sh) (Granlund and Montgomery,
// "Division by invariant integers using multiplication", 1994): with
// l = ceil(log2 d), sh = 31 + l and m = ceil(2^sh / d) < 2^32, the excess
// m d - 2^sh lies in [0, d) and so below 2^l, which keeps the quotient exact
// for every n < 2^31.
inline void fast_div_params(int32_t d, uint32_t* m, int32_t* sh) {
    if (d < 1) d = 1;
    int l = 0;
    while ((1LL << l) < d) l++;
    *m = (uint32_t)(((1ULL << (31 + l)) + (uint64_t)d - 1) / (uint64_t)d);
    *sh = 31 + l;
}

// Obstacle link mask of an interior node (extension: in-domain obstacles
// with half-way bounce-back): bit q set = the pull source x - c_q is an
// obstacle node, so f_q(x) = fs_opp(q)(x); kObstSolid = the node itself is
// an obstacle node (ImmersedSolid).
constexpr uint32_t kObstSolid = 1u << 31;

// A box of boundary nodes (one region of one block): nodes start ..
// start+count-1 of the block's boundary list, x fastest.
struct BndSeg {
    int32_t org[3];
    int32_t ext[3];
    int32_t start;
    int32_t count;
};
constexpr int kBndWG = kTX * kTY;  // boundary nodes per workgroup

// Ping-pong parity is kept per block in device memory: a step reads
// cur_in[b] (the buffer holding block b's latest fStage) and the block's
// first workgroup writes cur_out[b] (flipped unless the block is frozen),
// so independent instances can stop inside one captured launch sequence.

enum Mode : int {
    MODE_STEP = 0,        // gather, boundary, collide -> fs_next
    MODE_STEP_CHECK = 1,  // as STEP, plus macros (and residual partials
                          // when want_residual)
    MODE_MATERIALIZE = 2, // gather, boundary -> f (host layout), macros
    MODE_INIT_FEQ = 3,    // f = feq(macros) -> collide -> fs_next (and f
                          // to fhost when given)
    MODE_INIT_F = 4       // f from host-layout array -> collide -> fs_next
};

struct KParams {
    const BlockParams* blocks;
    int32_t nblocks;
    int32_t ntiles;
    const RegionProg* progs;
    void* fs_buf[2];        // the two fStage buffers
    void* per_buf[2];       // the two persist buffers
    const int32_t* cur_in;  // per-block buffer index of the latest fStage
    int32_t* cur_out;       // per-block index after this step (step modes)
    double* macro[4];       // MacroVars (current values, host layout)
    double* macro_copy[4];  // MacroVarsCopy (residual reference)
    // the macros the initial condition was formed from (SetInitialMacrosVars
    // as uploaded when PreDefinedInitialCondition was called); null: macro
    const double* macro0[4];
    double* fhost;      // materialize output / init input, host layout
    double* partials;   // per logical workgroup: kNAcc doubles
    float scale[kMaxQ];     // FP16 storage: g = h * scale[q]
    float inv_scale[kMaxQ];
    int32_t force_type;     // 0 none, 1 first order
    int32_t force_macros;   // *_Force velocity kernels
    int32_t fhost_in;       // MATERIALIZE: take f from fhost instead of streaming
    int32_t want_residual;  // accumulate the residual against MacroVarsCopy
    const uint32_t* obst;   // obstacle link masks per interior node, or null
    const BndSeg* segs;     // boundary segments, blocks in order
    int32_t fold;           // 0 none, 1 simple x faces, 2 every simple region of
                            // the planes z = 1 .. n-2, 3 every simple region
                            // (the bulk grid then spans every z plane) run
                            // inside the bulk kernel
    int32_t xcd_remap;      // bulk tiles dealt to XCDs in contiguous runs
    int32_t xgroup;         // bulk workgroup shape: 0 kTX x kTY tiles in order,
                            // 1 the same with each block's rows of tiles dealt
                            // to one XCD, 2 kTY x-tiles of one row (blocks with
                            // rowwg; the others as 0)
    int32_t fused;          // 1: the step is one whole-lattice launch
    int32_t h2;             // 1: FP16 steps use the two-node bulk kernel
    int32_t nbulk2;         // its workgroups
    int32_t nbulk;          // bulk workgroups (partials 0 .. nbulk-1)
    int32_t nbnd;           // boundary workgroups (partials nbulk ..)
    // Collision_BGKThermal4th (FP64 storage): per-node relaxation time and
    // the 4th-order equilibrium at the temperature temp (host layout, as
    // the macros)
    int32_t thermal;
    const double* temp;
    // the block of every logical workgroup, one table per grid (whole
    // lattice, bulk, bulk over every z plane, two-node FP16 bulk, boundary
    // at the fold level in force): one scalar load where a binary search
    // over the blocks' first workgroups took log2(nblocks) dependent ones
    const int32_t* wg_block_all;
    const int32_t* wg_block_bulk;
    const int32_t* wg_block_bulk3;
    const int32_t* wg_block_h2;
    const int32_t* wg_block_bnd;
};

struct HaloParams {
    const void* src_cur;    // buffer being exchanged (current step)
    const void* src_other;  // the other ping-pong buffer
    void* dst;
    int64_t src_base, dst_base;  // node (0,0,0) population 0 of each block
    int32_t src_sx, src_sz, src_sq;
    int32_t dst_sx, dst_sz, dst_sq;
    int32_t org_src[3], ext[3];
    int32_t Q;
    int32_t elem_bytes;
    // target element of slab cell (a, b, c), population q:
    // dst_base + qmap[q] * dst_sq + dst_org + a*dstep[0] + b*dstep[1] + c*dstep[2]
    // (a plain link has dstep = (1, dst_sx, dst_sz) and qmap[q] = q; a
    // rotated one turns both)
    int64_t dst_org;
    int64_t dstep[3];
    int32_t qmap[kMaxQ];
    // per slab cell (or null = all 0): bit 0 = read src_other; bit 1 = copy
    // every population (another connection reads the target cell)
    const uint8_t* other_flag;
    // populations (source index) that stream from the target halo into the
    // target block: the only ones a cell without bit 1 needs
    uint32_t inward;
    // per slab cell (or null): target populations a later connection of the
    // same merged launch writes at this cell's target, which this one
    // leaves to it
    const uint32_t* sup;
};

}  // namespace mplbhip
