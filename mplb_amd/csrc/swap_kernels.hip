// swap_kernels.hip -- the swap scheme (Scheme_StreamCollision_Swap) on gfx950.
//
// SwapStreamCollision (Src/evolution.cpp:195-252) keeps a single lattice,
// f, with halos, and updates it in place: macros -> BGK collision -> body
// force -> halo exchange of f -> local swap -> swap stream -> boundary
// conditions.  The velocity order puts opposites half a set apart (opp(i) =
// i + 9 for i = 1..9 in D3Q19), which the swaps rely on (J. Latt, "How to
// implement your DdQq dynamics with only q variables per node").  Only
// D3Q19 has swap kernels in the reference (scheme_kernel.inc:957-1398), and
// this path keeps FP64 storage: the operations are the reference's, in its
// order, so the results match the CPU oracle bit for bit.
//
// Each kernel covers one of the reference's loops, so each sees the
// previous one complete:
//   k_swap_collide  macros, collision in place, body force (every node)
//   (halo copies)   k_halo on f, every population
//   k_swap_local    KerLocalSwap3D: per node, in place
//   k_swap_stream   KerSwapStream3D: node x exchanges f_{i+9}(x) with
//                   f_i(x + c_i); no two exchanges touch the same value,
//                   so every node runs at once
//   k_swap_bc       the boundary applications of each boundary node in
//                   definition order (EQMDiffuseRefl, FDPeriodic)
//   k_swap_measure  macros, residual and mass partials, f for the host
#include <hip/hip_runtime.h>

#include "../../include/mplbhip.h"
#include "lattice_kernels.hpp"

namespace mplbhip {

namespace {

constexpr int Q = 19, HALF = 9;
// sqrt(3) rounded to double, as the reference's CS (model.cpp:67-68)
constexpr double kCS3 = 1.7320508075688772935274463415059;

__device__ __forceinline__ int cv(int q, int a) {
    constexpr int t[19][3] = {
        {0, 0, 0},  {-1, 0, 0},  {0, -1, 0}, {0, 0, -1}, {-1, -1, 0},
        {-1, 1, 0}, {-1, 0, -1}, {-1, 0, 1}, {0, -1, -1}, {0, -1, 1},
        {1, 0, 0},  {0, 1, 0},   {0, 0, 1},  {1, 1, 0},  {1, -1, 0},
        {1, 0, 1},  {1, 0, -1},  {0, 1, 1},  {0, 1, -1}};
    return t[q][a];
}
__device__ __forceinline__ double wq(int q) {
    return q == 0 ? 1.0 / 3.0 : ((q <= 3 || (q >= 10 && q <= 12)) ? 1.0 / 18.0 : 1.0 / 36.0);
}
__device__ __forceinline__ int oppq(int q) { return q == 0 ? 0 : (q <= HALF ? q + HALF : q - HALF); }

// CalcBGKFeq, T = 1, polyOrder = 2 (model_host_device.h:58-83), operation
// for operation
__device__ __forceinline__ double feq3(int q, double rho, double u, double v, double w) {
#pragma clang fp contract(off)
    const double cu = kCS3 * double(cv(q, 0)) * u + kCS3 * double(cv(q, 1)) * v +
                      kCS3 * double(cv(q, 2)) * w;
    const double u2 = u * u + v * v + w * w;
    const double res = 1.0 + cu + 0.5 * (cu * cu - u2);
    return wq(q) * rho * res;
}

// KerCutCellEQMDiffuseRefl3D (boundary_kernel.inc:1184-1276) in place
__device__ void eqm3(double* f, const BCOp& op) {
#pragma clang fp contract(off)
    const double u = op.vars[0], v = op.vars[1], w = op.vars[2];
    double rhoIn = 0, rhoPar = 0, dRho = 0;
    for (int q = 0; q < Q; q++) {
        const double cu = kCS3 * double(cv(q, 0)) * u + kCS3 * double(cv(q, 1)) * v +
                          kCS3 * double(cv(q, 2)) * w;
        if (op.m_in >> q & 1u) rhoIn += f[q];
        else if (op.m_out >> q & 1u) dRho += (2.0 * wq(q)) * cu;
        else if (op.m_par >> q & 1u) rhoPar += feq3(q, 1.0, u, v, w);
    }
    const double rhoW = 2.0 * rhoIn / (1.0 - dRho - rhoPar);
    for (int q = 0; q < Q; q++)
        if (op.m_par >> q & 1u) f[q] = feq3(q, rhoW, u, v, w);
    for (int q = 0; q < Q; q++)
        if (op.m_out >> q & 1u) {
            const double cu = kCS3 * double(cv(q, 0)) * u + kCS3 * double(cv(q, 1)) * v +
                              kCS3 * double(cv(q, 2)) * w;
            f[q] = f[oppq(q)] + 2.0 * rhoW * wq(q) * cu;
        }
}

// geometry code -> face normals (10 IP, 11 IM, 20 JP, 21 JM, 30 KP, 31 KM,
// two-digit codes concatenated for edges and corners, a trailing 0 for the
// inner kinds; flowfield_host_device.h:77-144); 0 for anything else
__device__ __forceinline__ int decode(int vg, int* ax, int* sg) {
    int nf, codes[3];
    if (vg >= 10 && vg < 100) {
        nf = 1;
        codes[0] = vg;
    } else if (vg >= 10000 && vg < 100000) {
        if (vg % 10) return 0;
        vg /= 10;
        nf = 2;
        codes[0] = vg / 100;
        codes[1] = vg % 100;
    } else if (vg >= 1000000 && vg < 10000000) {
        if (vg % 10) return 0;
        vg /= 10;
        nf = 3;
        codes[0] = vg / 10000;
        codes[1] = (vg / 100) % 100;
        codes[2] = vg % 100;
    } else {
        return 0;
    }
    for (int i = 0; i < nf; i++) {
        ax[i] = codes[i] / 10 - 1;
        sg[i] = codes[i] % 10 == 0 ? 1 : -1;
    }
    return nf;
}

__device__ __forceinline__ bool fluid_like(int vt) {
    return vt == VT_Fluid || vt == VT_MDPeriodic || vt == VT_VirtualBoundary;
}

// node of a whole-lattice tile (the tiling of k_lattice's SCOPE_ALL)
struct NodeRef {
    int b, x, y, z, wg;
    bool active;
};

__device__ __forceinline__ NodeRef tile_node(const KParams& kp) {
    NodeRef r;
    r.wg = (int)blockIdx.x;
    int lo = 0, hi = kp.nblocks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (kp.blocks[mid].tile_start <= r.wg) lo = mid; else hi = mid - 1;
    }
    r.b = lo;
    const BlockParams& B = kp.blocks[r.b];
    const int t = r.wg - B.tile_start;
    const int tx = t % B.tiles_x;
    const int t1 = t / B.tiles_x;
    const int ty = t1 % B.tiles_y;
    r.z = t1 / B.tiles_y;
    r.x = tx * kTX + (int)threadIdx.x;
    r.y = ty * kTY + (int)threadIdx.y;
    r.active = r.x < B.n[0] && r.y < B.n[1];
    return r;
}

__device__ __forceinline__ const RegionProg& region(const KParams& kp, const BlockParams& B,
                                                    int x, int y, int z) {
    const int r = (x == 0 ? 1 : (x == B.n[0] - 1 ? 2 : 0)) +
                  3 * (y == 0 ? 1 : (y == B.n[1] - 1 ? 2 : 0)) +
                  9 * (z == 0 ? 1 : (z == B.n[2] - 1 ? 2 : 0));
    return kp.progs[B.prog_base + r];
}

__device__ __forceinline__ double* fptr(const KParams& kp, const BlockParams& B) {
    return static_cast<double*>(kp.fs_buf[0]) + B.base;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace

// UpdateMacroVars3D (model_kernel.inc:570-759) then
// KerSwapCollideBGKIsothermal3D (:385-408, every node) then
// KerSwapCalcBodyForce1ST3D (:543-568, f += F at Fluid and MDPeriodic)
__global__ __launch_bounds__(kTX* kTY) void k_swap_collide(KParams kp) {
#pragma clang fp contract(off)
    const NodeRef r = tile_node(kp);
    if (!r.active) return;
    const BlockParams& B = kp.blocks[r.b];
    const RegionProg& P = region(kp, B, r.x, r.y, r.z);
    double* f = fptr(kp, B);
    const int64_t node = (int64_t)r.z * B.sz + (int64_t)r.y * B.sx + r.x;
    double fl[Q];
    for (int q = 0; q < Q; q++) fl[q] = f[(int64_t)q * B.sq + node];
    const bool forced = P.vt == VT_Fluid || P.vt == VT_MDPeriodic;
    double rho = 0;
    for (int q = 0; q < Q; q++) rho += fl[q];
    double m[3];
    for (int d = 0; d < 3; d++) {
        double u = 0;
        for (int q = 0; q < Q; q++) u += kCS3 * double(cv(q, d)) * fl[q];
        u /= rho;
        if (kp.force_macros && forced) u += (B.dt * B.g[d] / 2);
        m[d] = u;
    }
    for (int q = 0; q < Q; q++) {
        const double fe = feq3(q, rho, m[0], m[1], m[2]);
        double v = fe + B.one_minus_omega * (fl[q] - fe);
        if (kp.force_type == 1 && forced) {
            double cf = 0;
            for (int d = 0; d < 3; d++) cf += kCS3 * double(cv(q, d)) * B.g[d];
            const double F = wq(q) * rho * cf;
            v += F;
        }
        f[(int64_t)q * B.sq + node] = v;
    }
}

// KerLocalSwap3D (scheme_kernel.inc:959-1194): a fluid node exchanges f_i
// and f_{i+9}; a boundary node sets f_i = f_opp(i) for the velocities that
// point out of the block or along it on every normal of its geometry, not
// along all of them (those pairs are swapped twice by the reference's loop,
// an identity)
__global__ __launch_bounds__(kTX* kTY) void k_swap_local(KParams kp) {
    const NodeRef r = tile_node(kp);
    if (!r.active) return;
    const BlockParams& B = kp.blocks[r.b];
    const RegionProg& P = region(kp, B, r.x, r.y, r.z);
    if (P.vt == VT_ImmersedSolid) return;
    double* f = fptr(kp, B);
    const int64_t node = (int64_t)r.z * B.sz + (int64_t)r.y * B.sx + r.x;
    double fl[Q];
    for (int q = 0; q < Q; q++) fl[q] = f[(int64_t)q * B.sq + node];
    if (fluid_like(P.vt)) {
        for (int q = 1; q <= HALF; q++) {
            f[(int64_t)q * B.sq + node] = fl[q + HALF];
            f[(int64_t)(q + HALF) * B.sq + node] = fl[q];
        }
        return;
    }
    int ax[3], sg[3];
    const int nf = decode(P.vg, ax, sg);
    if (nf == 0) return;
    for (int q = 1; q < Q; q++) {
        bool allle = true, allzero = true;
        for (int a = 0; a < nf; a++) {
            const int cn = sg[a] * cv(q, ax[a]);
            allle = allle && cn <= 0;
            allzero = allzero && cn == 0;
        }
        if (allle && !allzero) f[(int64_t)q * B.sq + node] = fl[oppq(q)];
    }
}

// KerSwapStream3D (scheme_kernel.inc:1196-1398): node x exchanges
// f_{i+9}(x) with f_i(x + c_i), i = 1..9; a boundary node only where the
// opposite velocity does not point into the block along any normal of its
// geometry (sign * c_opp <= 0 on each)
__global__ __launch_bounds__(kTX* kTY) void k_swap_stream(KParams kp) {
    const NodeRef r = tile_node(kp);
    if (!r.active) return;
    const BlockParams& B = kp.blocks[r.b];
    const RegionProg& P = region(kp, B, r.x, r.y, r.z);
    if (P.vt == VT_ImmersedSolid) return;
    int ax[3] = {0, 0, 0}, sg[3] = {0, 0, 0}, nf = 0;
    if (!fluid_like(P.vt)) {
        nf = decode(P.vg, ax, sg);
        if (nf == 0) return;
    }
    double* f = fptr(kp, B);
    const int64_t node = (int64_t)r.z * B.sz + (int64_t)r.y * B.sx + r.x;
    for (int q = 1; q <= HALF; q++) {
        bool ok = true;
        for (int a = 0; a < nf; a++) ok = ok && sg[a] * cv(oppq(q), ax[a]) <= 0;
        if (!ok) continue;
        const int64_t nb = node + cv(q, 0) + (int64_t)cv(q, 1) * B.sx + (int64_t)cv(q, 2) * B.sz;
        double* a = f + (int64_t)(q + HALF) * B.sq + node;
        double* b = f + (int64_t)q * B.sq + nb;
        const double t = *a;
        *a = *b;
        *b = t;
    }
}

// ImplementBoundary3D (boundary.cpp:379-398) on the boundary segments:
// EQMDiffuseRefl in place, FDPeriodic (KerCutCellPeriodic3D,
// boundary_kernel.inc:1278-1666) copies every population from the cell
// outside the face, which no kernel of this stage writes
__global__ __launch_bounds__(kTX* kTY) void k_swap_bc(KParams kp) {
    const int wg = (int)blockIdx.x;
    int lo = 0, hi = kp.nblocks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (kp.blocks[mid].bnd_wg_start[0] <= wg) lo = mid; else hi = mid - 1;
    }
    const BlockParams& B = kp.blocks[lo];
    const int i = (wg - B.bnd_wg_start[0]) * kBndWG + (int)threadIdx.y * kTX + (int)threadIdx.x;
    int s = B.seg_start[0];
    const int send = s + B.seg_count[0];
    while (s < send && i >= kp.segs[s].start + kp.segs[s].count) s++;
    if (!(s < send && i >= kp.segs[s].start)) return;
    const BndSeg& S = kp.segs[s];
    int rr = i - S.start;
    const int x = S.org[0] + rr % S.ext[0];
    rr /= S.ext[0];
    const int y = S.org[1] + rr % S.ext[1];
    const int z = S.org[2] + rr / S.ext[1];
    const RegionProg& P = region(kp, B, x, y, z);
    if (P.nops == 0 || P.vt == VT_ImmersedSolid) return;
    double* f = fptr(kp, B);
    const int64_t node = (int64_t)z * B.sz + (int64_t)y * B.sx + x;
    double fl[Q];
    for (int q = 0; q < Q; q++) fl[q] = f[(int64_t)q * B.sq + node];
    for (int o = 0; o < P.nops; o++) {
        const BCOp& op = P.op[o];
        if (op.type == OP_EQM) {
            eqm3(fl, op);
        } else if (op.type == OP_FDPERIODIC) {
            const int64_t src = node + op.off[0] + (int64_t)op.off[1] * B.sx +
                                (int64_t)op.off[2] * B.sz;
            for (int q = 0; q < Q; q++) fl[q] = f[(int64_t)q * B.sq + src];
        }
    }
    for (int q = 0; q < Q; q++) f[(int64_t)q * B.sq + node] = fl[q];
}

// f^0 = feq of the initial macros at every non-solid node
// (KerInitialiseBGK2nd3D, model_kernel.inc:352-380), or f from a host-layout
// array (restart)
__global__ __launch_bounds__(kTX* kTY) void k_swap_init(KParams kp) {
    const NodeRef r = tile_node(kp);
    if (!r.active) return;
    const BlockParams& B = kp.blocks[r.b];
    const RegionProg& P = region(kp, B, r.x, r.y, r.z);
    double* f = fptr(kp, B);
    const int64_t node = (int64_t)r.z * B.sz + (int64_t)r.y * B.sx + r.x;
    const int64_t mnode = B.mbase + ((int64_t)r.z * B.n[1] + r.y) * B.n[0] + r.x;
    const int64_t nn = (int64_t)B.n[0] * B.n[1] * B.n[2];
    const int64_t h = B.fbase + ((int64_t)r.z * B.n[1] + r.y) * B.n[0] + r.x;
    const double* const* m0 = kp.macro0[0] ? kp.macro0 : kp.macro;
    for (int q = 0; q < Q; q++) {
        double v;
        if (kp.fhost_in) v = kp.fhost[h + q * nn];
        else if (P.vt == VT_ImmersedSolid) v = 0.0;
        else v = feq3(q, m0[0][mnode], m0[1][mnode], m0[2][mnode], m0[3][mnode]);
        f[(int64_t)q * B.sq + node] = v;
    }
}

// UpdateMacroVars3D on the current f (macros written, with the residual
// against MacroVarsCopy when want_residual), the FP64 mass and residual
// partials per workgroup, and f in host layout when fhost is given
__global__ __launch_bounds__(kTX* kTY) void k_swap_measure(KParams kp) {
#pragma clang fp contract(off)
    const NodeRef r = tile_node(kp);
    double acc[kNAcc];
    for (int a = 0; a < kNAcc; a++) acc[a] = 0.0;
    if (r.active) {
        const BlockParams& B = kp.blocks[r.b];
        const RegionProg& P = region(kp, B, r.x, r.y, r.z);
        const double* f = fptr(kp, B);
        const int64_t node = (int64_t)r.z * B.sz + (int64_t)r.y * B.sx + r.x;
        const int64_t mnode = B.mbase + ((int64_t)r.z * B.n[1] + r.y) * B.n[0] + r.x;
        double fl[Q];
        for (int q = 0; q < Q; q++) fl[q] = f[(int64_t)q * B.sq + node];
        if (P.vt != VT_ImmersedSolid) {
            const bool forced = P.vt == VT_Fluid || P.vt == VT_MDPeriodic;
            double rho = 0;
            for (int q = 0; q < Q; q++) rho += fl[q];
            double m[4] = {rho, 0, 0, 0};
            for (int d = 0; d < 3; d++) {
                double u = 0;
                for (int q = 0; q < Q; q++) u += kCS3 * double(cv(q, d)) * fl[q];
                u /= rho;
                if (kp.force_macros && forced) u += (B.dt * B.g[d] / 2);
                m[1 + d] = u;
            }
            for (int v = 0; v < 4; v++) {
                if (kp.want_residual) {
                    const double old = kp.macro_copy[v][mnode];
                    acc[v] = (m[v] - old) * (m[v] - old);
                    acc[4 + v] = m[v] * m[v];
                    kp.macro_copy[v][mnode] = m[v];
                }
                kp.macro[v][mnode] = m[v];
            }
        }
        double mass = 0;
        for (int q = 0; q < Q; q++) mass += fl[q];
        acc[8] = mass;
        if (kp.fhost) {
            const int64_t nn = (int64_t)B.n[0] * B.n[1] * B.n[2];
            const int64_t h = B.fbase + ((int64_t)r.z * B.n[1] + r.y) * B.n[0] + r.x;
            for (int q = 0; q < Q; q++) kp.fhost[h + q * nn] = fl[q];
        }
    }
    // the workgroup's partial sums (the slot layout of k_lattice's
    // whole-lattice launches, which launch_reduce adds per block)
    __shared__ double red[kTY][kNAcc];
    for (int a = 0; a < kNAcc; a++) {
        const double s = wave_sum_d(acc[a]);
        if (threadIdx.x == 0) red[threadIdx.y][a] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        for (int a = 0; a < kNAcc; a++) {
            double s = 0.0;
            for (int k = 0; k < kTY; k++) s += red[k][a];
            kp.partials[(int64_t)r.wg * kNAcc + a] = s;
        }
    }
}

// part: 0 collide, 1 local swap, 2 swap stream, 3 boundaries, 4 init,
// 5 measure
hipError_t launch_swap(int part, const KParams& kp, int nbnd, hipStream_t s) {
    const dim3 blk(kTX, kTY, 1), grd(kp.ntiles, 1, 1);
    switch (part) {
        case 0: hipLaunchKernelGGL(k_swap_collide, grd, blk, 0, s, kp); break;
        case 1: hipLaunchKernelGGL(k_swap_local, grd, blk, 0, s, kp); break;
        case 2: hipLaunchKernelGGL(k_swap_stream, grd, blk, 0, s, kp); break;
        case 3:
            if (nbnd > 0) hipLaunchKernelGGL(k_swap_bc, dim3(nbnd), blk, 0, s, kp);
            break;
        case 4: hipLaunchKernelGGL(k_swap_init, grd, blk, 0, s, kp); break;
        case 5: hipLaunchKernelGGL(k_swap_measure, grd, blk, 0, s, kp); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mplbhip
