// lattice_kernels.hip -- gfx950 kernels of the MPLB lattice update.
//
// One fused pull kernel per step: a thread owns one node, gathers the
// post-collision populations of the previous step from its neighbours
// (stream, scheme_kernel.inc:53-221 / 1400-2052), applies the boundary
// program of its region (boundary_kernel.inc), computes the macroscopic
// variables (model_kernel.inc:57-186 / 570-759) and collides
// (model_kernel.inc:221-265 / 410-455), writing the post-collision state
// to the other ping-pong buffer.  The state kept in HBM is therefore fStage
// (the reference's post-collision field) plus, for the few boundary regions
// whose unstreamed populations must survive a step, a compact persist
// buffer holding their post-boundary f.
//
// FP64 storage reproduces the reference arithmetic operation for operation
// (contraction off, same summation order, IEEE division), so the parity
// path matches the CPU oracle bit for bit.  FP32/FP16 storage keeps
// g = f - w_i and does the bulk update in FP32 in deviation form.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "../../include/mplbhip.h"
#include "lattice_kernels.hpp"

namespace mplbhip {

// sqrt(3) rounded to double: equals the host's sqrt(3.0) (model.cpp:67-68)
constexpr double kCS = 1.7320508075688772935274463415059;

// The lattice key D of every template below: 2 = D2Q9, 3 = D3Q19,
// 15 = D3Q15 (Src/model.cpp:133-150); ldim(D) is its space dimension and
// lq(D) its number of velocities.
__host__ __device__ constexpr int ldim(int D) { return D == 2 ? 2 : 3; }
__host__ __device__ constexpr int lq(int D) { return D == 2 ? 9 : (D == 15 ? 15 : 19); }

__host__ __device__ constexpr int cvel(int D, int q, int a) {
    constexpr int t2[9][2] = {{0, 0},  {-1, 1}, {-1, 0}, {-1, -1}, {0, -1},
                              {1, -1}, {1, 0},  {1, 1},  {0, 1}};
    constexpr int t3[19][3] = {
        {0, 0, 0},  {-1, 0, 0},  {0, -1, 0}, {0, 0, -1}, {-1, -1, 0},
        {-1, 1, 0}, {-1, 0, -1}, {-1, 0, 1}, {0, -1, -1}, {0, -1, 1},
        {1, 0, 0},  {0, 1, 0},   {0, 0, 1},  {1, 1, 0},  {1, -1, 0},
        {1, 0, 1},  {1, 0, -1},  {0, 1, 1},  {0, 1, -1}};
    constexpr int t15[15][3] = {
        {0, 0, 0},    {-1, 0, 0},  {0, -1, 0},  {0, 0, -1}, {-1, -1, -1},
        {-1, -1, 1},  {-1, 1, -1}, {-1, 1, 1},  {1, 0, 0},  {0, 1, 0},
        {0, 0, 1},    {1, 1, 1},   {1, 1, -1},  {1, -1, 1}, {1, -1, -1}};
    return D == 2 ? (a < 2 ? t2[q][a] : 0) : (D == 15 ? t15[q][a] : t3[q][a]);
}
__host__ __device__ constexpr double wgt(int D, int q) {
    constexpr double w2[9] = {4.0 / 9.0,  1.0 / 36.0, 1.0 / 9.0,
                              1.0 / 36.0, 1.0 / 9.0,  1.0 / 36.0,
                              1.0 / 9.0,  1.0 / 36.0, 1.0 / 9.0};
    constexpr double w3[19] = {
        1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0,
        1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
        1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0,
        1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    // 2/9, 1/9, 1/72 as the reference writes them: 2 / ((Real)9) ...
    constexpr double w15[15] = {
        2.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 72.0,
        1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 9.0,  1.0 / 9.0,
        1.0 / 9.0,  1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
    return D == 2 ? w2[q] : (D == 15 ? w15[q] : w3[q]);
}
__host__ __device__ constexpr int opp(int D, int q) {
    constexpr int o2[9] = {0, 5, 6, 7, 8, 1, 2, 3, 4};
    constexpr int o3[19] = {0, 10, 11, 12, 13, 14, 15, 16, 17, 18,
                            1, 2,  3,  4,  5,  6,  7,  8,  9};
    constexpr int o15[15] = {0, 8, 9, 10, 11, 12, 13, 14, 1, 2, 3, 4, 5, 6, 7};
    return D == 2 ? o2[q] : (D == 15 ? o15[q] : o3[q]);
}

// ---- storage traits -------------------------------------------------------
template <int ST> struct Store;
template <> struct Store<MPLBHIP_STORE_F64> { using E = double; using T = double; };
template <> struct Store<MPLBHIP_STORE_F32> { using E = float; using T = float; };
template <> struct Store<MPLBHIP_STORE_F16> { using E = __half; using T = float; };

// a stored FP16 value times its population scale: one mixed-precision
// multiply-add (v_fma_mix) instead of a conversion and a multiply
__device__ __forceinline__ float h2f_scaled(__half h, float sc) {
    return __builtin_fmaf(__half2float(h), sc, 0.0f);
}

template <int ST>
__device__ __forceinline__ typename Store<ST>::T ld(const typename Store<ST>::E* p,
                                                    int64_t i, float sc) {
    if constexpr (ST == MPLBHIP_STORE_F16) return h2f_scaled(p[i], sc);
    else return p[i];
}
template <int ST>
__device__ __forceinline__ void st(typename Store<ST>::E* p, int64_t i,
                                   typename Store<ST>::T v, float isc) {
    if constexpr (ST == MPLBHIP_STORE_F16) p[i] = __float2half_rn(v * isc);
    else p[i] = v;
}
// non-temporal store for the streamed post-collision state
template <int ST>
__device__ __forceinline__ void st_nt(typename Store<ST>::E* p, int64_t i,
                                      typename Store<ST>::T v, float isc) {
    if constexpr (ST == MPLBHIP_STORE_F16) {
        p[i] = __float2half_rn(v * isc);
    } else {
        __builtin_nontemporal_store(v, p + i);
    }
}

// ---- reference arithmetic (any T; exact for double) -------------------------
// CalcBGKFeq, T = 1, polyOrder = 2 (model_host_device.h:34-83).  The
// (T-1)(c2-D) term is an exact zero and is left out.
template <int D, typename T>
__device__ __forceinline__ T feq_ref(int q, T rho, T u, T v, T w) {
#pragma clang fp contract(off)
    const T cs = T(kCS);
    T cu;
    T u2;
    if constexpr (D == 2) {
        cu = cs * T(cvel(D, q, 0)) * u + cs * T(cvel(D, q, 1)) * v;
        u2 = u * u + v * v;
    } else {
        cu = cs * T(cvel(D, q, 0)) * u + cs * T(cvel(D, q, 1)) * v +
             cs * T(cvel(D, q, 2)) * w;
        u2 = u * u + v * v + w * w;
    }
    const T cu2 = cu * cu;
    const T res = T(1.0) + cu + T(0.5) * (cu2 - u2);
    return T(wgt(D, q)) * rho * res;
}

// CalcBGKFeq at temperature Tm with polyOrder 4 (model_host_device.h:34-83),
// every term in the reference's order; LATTDIM enters as an int
template <int D>
__device__ __forceinline__ double feq4_ref(int q, double rho, double u, double v, double w,
                                           double Tm) {
#pragma clang fp contract(off)
    constexpr int LD = ldim(D);
    const double cs = kCS;
    double cu, c2, u2;
    if constexpr (LD == 2) {
        cu = (cs * double(cvel(D, q, 0)) * u + cs * double(cvel(D, q, 1)) * v);
        c2 = (cs * double(cvel(D, q, 0)) * cs * double(cvel(D, q, 0)) +
              cs * double(cvel(D, q, 1)) * cs * double(cvel(D, q, 1)));
        u2 = u * u + v * v;
    } else {
        cu = (cs * double(cvel(D, q, 0)) * u + cs * double(cvel(D, q, 1)) * v +
              cs * double(cvel(D, q, 2)) * w);
        c2 = (cs * double(cvel(D, q, 0)) * cs * double(cvel(D, q, 0)) +
              cs * double(cvel(D, q, 1)) * cs * double(cvel(D, q, 1)) +
              cs * double(cvel(D, q, 2)) * cs * double(cvel(D, q, 2)));
        u2 = u * u + v * v + w * w;
    }
    const double cu2 = cu * cu;
    double res = 1.0 + cu + 0.5 * (cu2 - u2 + (Tm - 1.0) * (c2 - LD));
    res = res + cu * (cu2 - 3.0 * u2 + 3.0 * (Tm - 1.0) * (c2 - LD - 2.0)) / 6.0;
    res = res + (cu2 * cu2 - 6.0 * cu2 * u2 + 3.0 * u2 * u2) / 24.0 +
          (Tm - 1.0) * ((c2 - (LD + 2)) * (cu2 - u2) - 2.0 * cu2) / 4.0 +
          (Tm - 1.0) * (Tm - 1.0) * (c2 * c2 - 2.0 * (LD + 2) * c2 + LD * (LD + 2)) / 8.0;
    return wgt(D, q) * rho * res;
}

template <int D, typename T>
__device__ __forceinline__ T cu_wall(int q, T u, T v, T w) {
#pragma clang fp contract(off)
    const T cs = T(kCS);
    const T cx = cs * T(cvel(D, q, 0)), cy = cs * T(cvel(D, q, 1));
    if constexpr (D == 2) {
        return cx * u + cy * v;
    } else {
        const T cz = cs * T(cvel(D, q, 2));
        return cx * u + cy * v + cz * w;
    }
}

// KerCutCellEQMDiffuseRefl[3D] (boundary_kernel.inc:534-613, 1184-1276)
template <int D, typename T>
__device__ void apply_eqm(T* f, const BCOp& op) {
#pragma clang fp contract(off)
    constexpr int Q = lq(D);
    const T u = T(op.vars[0]), v = T(op.vars[1]), w = T(op.vars[2]);
    T rhoIn = T(0), rhoPar = T(0), dRho = T(0);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        if (op.m_in >> q & 1u) {
            rhoIn += f[q];
        } else if (op.m_out >> q & 1u) {
            dRho += (T(2) * T(wgt(D, q))) * cu_wall<D, T>(q, u, v, w);
        } else if (op.m_par >> q & 1u) {
            rhoPar += feq_ref<D, T>(q, T(1), u, v, w);
        }
    }
    const T rhoW = T(2) * rhoIn / (T(1) - dRho - rhoPar);
#pragma unroll
    for (int q = 0; q < Q; q++)
        if (op.m_par >> q & 1u) f[q] = feq_ref<D, T>(q, rhoW, u, v, w);
#pragma unroll
    for (int q = 0; q < Q; q++)
        if (op.m_out >> q & 1u)
            f[q] = f[opp(D, q)] +
                   T(2) * rhoW * T(wgt(D, q)) * cu_wall<D, T>(q, u, v, w);
}

// apply_eqm on deviations g = f - w (reduced storage, FP32 arithmetic),
// with the terms that depend on the wall velocity only computed by the
// host (BCOp::eq_*): rho_in = sum_in w + sum_in g, rho_w = 2 rho_in /
// (1 - dRho - rhoPar), parallel g = rho_w w res - w, outgoing g = g_opp +
// 2 rho_w w c.u (the w of a population and its opposite are equal)
template <int D, typename OP>
__device__ __forceinline__ void apply_eqm_dev(float* g, const OP& op) {
    constexpr int Q = lq(D);
    const uint32_t m_in = op.m_in, m_out = op.m_out, m_par = op.m_par;
    // the incoming sum as one multiply-add per population with a uniform
    // 0/1 factor (a scalar select), not an add and a vector select
    float rin = op.eq_win;
#pragma unroll
    for (int q = 0; q < Q; q++) rin = __builtin_fmaf(g[q], (m_in >> q & 1u) ? 1.f : 0.f, rin);
    const float rw = rin * op.eq_k;
#pragma unroll
    for (int q = 0; q < Q; q++)
        if (m_par >> q & 1u) g[q] = __builtin_fmaf(rw, op.eq_c[q], -float(wgt(D, q)));
#pragma unroll
    for (int q = 0; q < Q; q++)
        if (m_out >> q & 1u) g[q] = __builtin_fmaf(rw, op.eq_c[q], g[opp(D, q)]);
}

// rescale part of KerCutCellExtrapolPressure1ST[3D]
// (boundary_kernel.inc:404-488, 894-1182); f already holds the copied
// populations.  The loop stops before the last population (:484, :1178).
template <int D, typename T>
__device__ void extrapol_rescale(T* f, double rho_given) {
#pragma clang fp contract(off)
    constexpr int Q = lq(D);
    T rho = T(0);
#pragma unroll
    for (int q = 0; q < Q; q++) rho += f[q];
    const T ratio = T(rho_given) / rho;
#pragma unroll
    for (int q = 0; q < Q - 1; q++) f[q] *= ratio;
}

// ---- node addressing --------------------------------------------------------
template <int D>
__device__ __forceinline__ int32_t pull_delta(const BlockParams& B, int q) {
    return cvel(D, q, 0) + cvel(D, q, 1) * B.sx + cvel(D, q, 2) * B.sz;
}

// element q*sq + node - delta of a block array: the uniform part forms a
// scalar base, the per-thread part a 32-bit byte offset (saddr addressing)
template <typename E>
__device__ __forceinline__ const E* at_u(const E* base, int64_t uniform, uint32_t node) {
    return reinterpret_cast<const E*>(reinterpret_cast<const char*>(base + uniform) +
                                      node * (uint32_t)sizeof(E));
}
template <typename E>
__device__ __forceinline__ E* at_u(E* base, int64_t uniform, uint32_t node) {
    return reinterpret_cast<E*>(reinterpret_cast<char*>(base + uniform) +
                                node * (uint32_t)sizeof(E));
}

// Raw buffer accesses (address = resource base + lane offset + scalar
// offset): the per-population part of an address is a scalar operand of
// the instruction, so a population costs no vector address arithmetic.
// Word 3 0x00020000 is the raw (untyped) buffer format; the record count
// 0xFFFFFFFF leaves every offset below 4 GiB in range.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(uint64_t base) {
    return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void*>(base), 0, -1, 0x00020000);
}
template <typename E>
__device__ __forceinline__ E buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(E) == 2)
        return __ushort_as_half(__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
    else if constexpr (sizeof(E) == 4)
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    else
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
// nontemporal store (cache policy bit 1: nt)
template <typename E>
__device__ __forceinline__ void buf_st_nt(E v, __amdgpu_buffer_rsrc_t r, uint32_t voff,
                                          uint32_t soff) {
    if constexpr (sizeof(E) == 2)
        __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(v), r, voff, soff, 2);
    else if constexpr (sizeof(E) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 2);
    else
        __builtin_amdgcn_raw_buffer_store_b64(
            __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r,
            voff, soff, 2);
}

__device__ __forceinline__ int region_code(int x, int n) {
    return x == 0 ? 1 : (x == n - 1 ? 2 : 0);
}

template <int D, typename PR = RegionProg>
__device__ __forceinline__ int64_t persist_slot(const PR& P, int x, int y, int z) {
    const int l0 = P.ext[0] > 1 ? x - 1 : 0;
    const int l1 = P.ext[1] > 1 ? y - 1 : 0;
    const int l2 = (D != 2 && P.ext[2] > 1) ? z - 1 : 0;
    return (int64_t)P.persist + l0 + (int64_t)P.ext[0] * (l1 + (int64_t)P.ext[1] * l2);
}

// obstacle link mask of interior node mnode (0 when the block has none)
__device__ __forceinline__ uint32_t obst_mask(const KParams& kp, const BlockParams& B,
                                              int64_t mnode) {
    return (kp.obst && B.has_obst) ? kp.obst[mnode] : 0u;
}

// residual terms of a node whose macros do not change (ImmersedSolid):
// (m - m_copy)^2 and m^2, with the copy refreshed, as the whole-range sums
// of CalcResidualError include them (flowfield_wrapper.cpp:31-89)
template <int D, int MODE>
__device__ __forceinline__ void solid_residual(const KParams& kp, int64_t mnode, double* acc) {
    if constexpr (MODE == MODE_STEP_CHECK || MODE == MODE_MATERIALIZE) {
        if (kp.want_residual) {
#pragma unroll
            for (int v = 0; v < ldim(D) + 1; v++) {
                const double m = kp.macro[v][mnode];
                const double old = kp.macro_copy[v][mnode];
                acc[v] = (m - old) * (m - old);
                acc[4 + v] = m * m;
                kp.macro_copy[v][mnode] = m;
            }
        }
    }
}

// full-value state of a node at the start of the boundary stage: streamed
// populations from the neighbours, unstreamed ones from the persist buffer
// PR: RegionProg, or the constant-address-space RegionProg of a
// wave-uniform program (scalar loads)
template <int D, int ST, typename PR = RegionProg>
__device__ void gather_masked(const KParams& kp, const BlockParams& B,
                              const PR& P, int x, int y, int z,
                              typename Store<ST>::T* f, const void* fs_prev,
                              const void* per_prev, uint32_t om) {
    using T = typename Store<ST>::T;
    using E = typename Store<ST>::E;
    constexpr int Q = lq(D);
    const E* prev = static_cast<const E*>(fs_prev) + B.base;
    const int32_t node = z * B.sz + y * B.sx + x;
    const T* per = static_cast<const T*>(per_prev);
    int64_t slot = P.persist >= 0 ? persist_slot<D, PR>(P, x, y, z) : 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        T v;
        if (P.pull >> q & 1u) {
            // an obstacle source bounces back the node's own opposite value
            if (om >> q & 1u)
                v = ld<ST>(prev, (int64_t)opp(D, q) * B.sq + node, kp.scale[opp(D, q)]);
            else
                v = ld<ST>(prev, (int64_t)q * B.sq + node - pull_delta<D>(B, q),
                           kp.scale[q]);
            if constexpr (ST != MPLBHIP_STORE_F64) v += T(wgt(D, q));
        } else if (P.persist >= 0) {
            v = per[slot * Q + q];
        } else {
            v = T(0);
        }
        f[q] = v;
    }
}

// one population of a neighbour, as the extrapolation reads it: after the
// neighbour's boundary applications that precede it in definition order
template <int D, int ST>
__device__ void neighbour_values(const KParams& kp, const BlockParams& B,
                                 const BCOp& op, int x, int y, int z,
                                 typename Store<ST>::T* f, const void* fs_prev,
                                 const void* per_prev) {
    using T = typename Store<ST>::T;
    constexpr int Q = lq(D);
    const RegionProg& N = kp.progs[B.prog_base + op.nb_region];
    const int nx = x + op.off[0], ny = y + op.off[1], nz = z + op.off[2];
    const int64_t nm = B.mbase + ((int64_t)nz * B.n[1] + ny) * B.n[0] + nx;
    T g[Q];
    gather_masked<D, ST>(kp, B, N, nx, ny, nz, g, fs_prev, per_prev, obst_mask(kp, B, nm));
    for (int o = 0; o < op.nb_nops; o++) {
        if (N.op[o].type == OP_EQM) apply_eqm<D, T>(g, N.op[o]);
    }
#pragma unroll
    for (int q = 0; q < Q; q++)
        if (op.m_ext >> q & 1u) f[q] = g[q];
}

template <int D, int ST, typename PR = RegionProg>
__device__ void boundary_program(const KParams& kp, const BlockParams& B,
                                 const PR& P, int x, int y, int z,
                                 typename Store<ST>::T* f, const void* fs_prev,
                                 const void* per_prev) {
    using T = typename Store<ST>::T;
    for (int o = 0; o < P.nops; o++) {
        BCOp op;
        __builtin_memcpy(&op, (const BCOp*)&P.op[o], sizeof op);
        if (op.type == OP_EQM) {
            apply_eqm<D, T>(f, op);
        } else if (op.type == OP_EXTRAPOL) {
            if (op.m_ext) neighbour_values<D, ST>(kp, B, op, x, y, z, f, fs_prev, per_prev);
            extrapol_rescale<D, T>(f, op.vars[0]);
        }
    }
}

// ---- workgroup reduction of the residual accumulators ------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ void reduce_store(double* acc, double* out) {
    __shared__ double red[kTY][kNAcc];
    const int lane = threadIdx.x, wv = threadIdx.y;
#pragma unroll
    for (int a = 0; a < kNAcc; a++) {
        const double s = wave_sum(acc[a]);
        if (lane == 0) red[wv][a] = s;
    }
    __syncthreads();
    if (lane == 0 && wv == 0) {
        for (int a = 0; a < kNAcc; a++) {
            double s = 0.0;
            for (int k = 0; k < kTY; k++) s += red[k][a];
            out[a] = s;
        }
    }
}

// logical workgroup -> XCD-contiguous chunk: hardware deals workgroups to
// the 8 XCDs round robin; this hands each XCD a contiguous run of tiles so
// that neighbouring tiles (and their halo re-reads) share an L2.
__device__ __forceinline__ int logical_wg(int wg, int n) {
    if ((n & 7) != 0) return wg;
    return (wg & 7) * (n >> 3) + (wg >> 3);
}

// workgroup -> tile such that the tx tiles of one row go to the same XCD
// (workgroups b, b+8, b+16, ... share an XCD) while consecutive rows are
// spread over the 8 XCDs; a bijection when the grid is a multiple of 8*tx
__device__ __forceinline__ int fast_div(int n, uint32_t m, int sh) {
    return (int)(((uint64_t)(uint32_t)n * m) >> sh);
}

// (the division by tx as the block's multiply-high-and-shift pair m, sh)
__device__ __forceinline__ int row_group_wg(int b, int tx, uint32_t m, int sh) {
    const int j = b >> 3;
    const int jq = fast_div(j, m, sh);
    return (jq * 8 + (b & 7)) * tx + (j - jq * tx);
}

// Block parameters and parities are never written while a step kernel
// runs.  Read through the constant address space, a wave-uniform index
// turns their loads into scalar loads: one batched round trip per wave
// instead of a chain of dependent vector loads ahead of the population
// loads.
#define MPLB_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const MPLB_CONST T* const_view(const T* p) {
    return (const MPLB_CONST T*)(p);
}

__device__ __forceinline__ BlockParams load_block(const KParams& kp, int b) {
    BlockParams B;
    __builtin_memcpy(&B, (const BlockParams*)&const_view(kp.blocks)[b], sizeof B);
    return B;
}

// the pointer held by the first active lane, as a wave-uniform value
template <typename T>
__device__ __forceinline__ const T* first_lane_ptr(const T* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return reinterpret_cast<const T*>(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int load_parity(const KParams& kp, int b) {
    return const_view(kp.cur_in)[b];
}

// block of logical workgroup wg of a grid, from that grid's table (wg is
// wave-uniform: a scalar load); a one-block case reads nothing
__device__ __forceinline__ int wg_block(const KParams& kp, const int32_t* table, int wg) {
    return kp.nblocks == 1 ? 0 : const_view(table)[wg];
}

// macros (+ residual partials), then the MATERIALIZE output or the BGK
// collision and the store of fStage.  f holds full values when `full`
// (boundary path), else (reduced storage) deviations g = f - w.
// BUF: the stores as buffer stores (the bulk kernel; next is wave-uniform).
template <int D, int ST, int MODE, bool BUF = false>
__device__ __forceinline__ void finish_node(const KParams& kp, const BlockParams& B,
                                            typename Store<ST>::T* f, bool full,
                                            bool force_node, int32_t node, int64_t mnode,
                                            typename Store<ST>::E* next, int x, int y,
                                            int z, double* acc) {
    using T = typename Store<ST>::T;
    using E = typename Store<ST>::E;
    constexpr int Q = lq(D);
    constexpr bool EXACT = ST == MPLBHIP_STORE_F64;
    (void)full;
    if constexpr (EXACT) {
#pragma clang fp contract(off)
        // UpdateMacroVars: density then each velocity component
        double rho = 0;
#pragma unroll
        for (int q = 0; q < Q; q++) rho += f[q];
        double uvw[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int d = 0; d < ldim(D); d++) {
            double u = 0;
#pragma unroll
            for (int q = 0; q < Q; q++) u += kCS * double(cvel(D, q, d)) * f[q];
            u /= rho;
            if (kp.force_macros && force_node) u += (B.dt * B.g[d] / 2);
            uvw[d] = u;
        }
        if constexpr (MODE == MODE_MATERIALIZE || MODE == MODE_STEP_CHECK) {
            const double m[4] = {rho, uvw[0], uvw[1], uvw[2]};
#pragma unroll
            for (int v = 0; v < ldim(D) + 1; v++) {
                if (kp.want_residual) {
                    const double old = kp.macro_copy[v][mnode];
                    acc[v] = (m[v] - old) * (m[v] - old);
                    acc[4 + v] = m[v] * m[v];
                    kp.macro_copy[v][mnode] = m[v];
                }
                kp.macro[v][mnode] = m[v];
            }
            double mass = 0;
#pragma unroll
            for (int q = 0; q < Q; q++) mass += f[q];
            acc[8] = mass;
        }
        if constexpr (MODE == MODE_MATERIALIZE) {
            if (kp.fhost && !kp.fhost_in) {
                const int64_t nn = (int64_t)B.n[0] * B.n[1] * B.n[2];
                const int64_t h = B.fbase + ((int64_t)z * B.n[1] + y) * B.n[0] + x;
#pragma unroll
                for (int q = 0; q < Q; q++) kp.fhost[h + q * nn] = f[q];
            }
        } else if (kp.thermal) {
            // KerCollideBGKThermal[3D] (model_kernel.inc:267-310, 457-502):
            // tau = tauRef / (rho sqrt(T)) per node, the 4th-order
            // equilibrium at T, and f - w (f - feq) (+ tau w F)
            const double Tm = kp.temp[mnode];
            const double tau = B.tau / (rho * sqrt(Tm));
            const double om = B.dt / (tau + 0.5 * B.dt);
            const bool addf = force_node && kp.force_type == 1;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const double fe = feq4_ref<D>(q, rho, uvw[0], uvw[1], uvw[2], Tm);
                double fsq = f[q] - om * (f[q] - fe);
                if (addf) {
                    double cf = 0;
#pragma unroll
                    for (int d = 0; d < ldim(D); d++)
                        cf += kCS * double(cvel(D, q, d)) * B.g[d];
                    const double F = wgt(D, q) * rho * cf;
                    fsq = fsq + tau * om * F;
                }
                *at_u(next, (int64_t)q * B.sq, (uint32_t)node) = fsq;
            }
        } else {
            // BGK collision with the optional first-order force source
            const double tw = B.tau_omega, om1 = B.one_minus_omega;
            const bool addf = force_node && kp.force_type == 1;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const double fe = feq_ref<D, double>(q, rho, uvw[0], uvw[1], uvw[2]);
                double fsq;
                if (addf) {
                    double cf = 0;
#pragma unroll
                    for (int d = 0; d < ldim(D); d++)
                        cf += kCS * double(cvel(D, q, d)) * B.g[d];
                    const double F = wgt(D, q) * rho * cf;
                    fsq = fe + om1 * (f[q] - fe) + tw * F;
                } else {
                    fsq = fe + om1 * (f[q] - fe);
                }
                *at_u(next, (int64_t)q * B.sq, (uint32_t)node) = fsq;
            }
        }
    } else {
        // deviation form g = f - w, FP32 arithmetic
        if (full) {
#pragma unroll
            for (int q = 0; q < Q; q++) f[q] -= T(wgt(D, q));
        }
        float drho = 0.f, j[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < Q; q++) {
            drho += f[q];
#pragma unroll
            for (int d = 0; d < ldim(D); d++) {
                if (cvel(D, q, d) > 0) j[d] += f[q];
                if (cvel(D, q, d) < 0) j[d] -= f[q];
            }
        }
        const float rho = 1.f + drho;
        const float inv = __builtin_amdgcn_rcpf(rho);  // v_rcp_f32, 1 ulp
        float ul[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int d = 0; d < ldim(D); d++) {
            ul[d] = j[d] * inv;
            if (kp.force_macros && force_node) ul[d] += B.fhalf_dtg[d];
        }
        const float usq = ul[0] * ul[0] + ul[1] * ul[1] + ul[2] * ul[2];
        if constexpr (MODE == MODE_MATERIALIZE || MODE == MODE_STEP_CHECK) {
            double dr = 0.0;
#pragma unroll
            for (int q = 0; q < Q; q++) dr += double(f[q]);
            const double m[4] = {1.0 + dr, kCS * double(ul[0]), kCS * double(ul[1]),
                                 kCS * double(ul[2])};
#pragma unroll
            for (int v = 0; v < ldim(D) + 1; v++) {
                if (kp.want_residual) {
                    const double old = kp.macro_copy[v][mnode];
                    acc[v] = (m[v] - old) * (m[v] - old);
                    acc[4 + v] = m[v] * m[v];
                    kp.macro_copy[v][mnode] = m[v];
                }
                kp.macro[v][mnode] = m[v];
            }
            acc[8] = 1.0 + dr;
        }
        if constexpr (MODE == MODE_MATERIALIZE) {
            if (kp.fhost && !kp.fhost_in) {
                const int64_t nn = (int64_t)B.n[0] * B.n[1] * B.n[2];
                const int64_t h = B.fbase + ((int64_t)z * B.n[1] + y) * B.n[0] + x;
#pragma unroll
                for (int q = 0; q < Q; q++)
                    kp.fhost[h + q * nn] = double(f[q]) + wgt(D, q);
            }
        } else {
            const float om = B.fomega, om1 = B.fone_minus_omega;
            const bool addf = force_node && kp.force_type == 1;
            const __amdgpu_buffer_rsrc_t rput = buf_rsrc(reinterpret_cast<uint64_t>(next));
            // hn: the new value in storage units (h = g / s_q for FP16)
            auto store = [&](int q, float hn) {
                if constexpr (BUF) {
                    const uint32_t voff = (uint32_t)node * (uint32_t)sizeof(E);
                    if constexpr (ST == MPLBHIP_STORE_F16)
                        buf_st_nt<E>(__float2half_rn(hn), rput, voff, B.put_off[q]);
                    else buf_st_nt<E>(hn, rput, voff, B.put_off[q]);
                    return;
                }
                E* p = at_u(next, (int64_t)q * B.sq, (uint32_t)node);
                if constexpr (ST == MPLBHIP_STORE_F16)
                    __builtin_nontemporal_store(__half_as_ushort(__float2half_rn(hn)),
                                                reinterpret_cast<unsigned short*>(p));
                else __builtin_nontemporal_store(hn, p);
            };
            const float c0 = drho + rho * (-1.5f * usq);
            if constexpr (ST == MPLBHIP_STORE_F16) {
                // g' = (1 - om) g + om geq in storage units h = g / s_q
                // (s_q = w_q / kF16Scale).  Opposite velocities share the
                // even part of the equilibrium: geq(q), geq(opp q) = w_q (A
                // +- B) with A = drho + rho (4.5 cu^2 - 1.5 u^2), B = 3 rho
                // cu, so h' = m_q g + k (A +- B), m_q = (1 - om) / s_q, k =
                // om kF16Scale (w_q cancels); the force term changes sign too
                const float k = om * float(kF16Scale);
                auto m_of = [&](int q) { return om1 * float(kF16Scale / wgt(D, q)); };
                store(0, __builtin_fmaf(m_of(0), f[0], k * c0));
                const float ka = (k * 4.5f) * rho, kb = (k * 3.f) * rho, kc = k * c0;
#pragma unroll
                for (int q = 1; q < Q; q++) {
                    const int p = opp(D, q);
                    if (p < q) continue;
                    // c.u without a leading 0 + (which would not fold away)
                    float cu = 0.f;
                    bool any = false;
#pragma unroll
                    for (int d = 0; d < ldim(D); d++) {
                        if (cvel(D, q, d) > 0) { cu = any ? cu + ul[d] : ul[d]; any = true; }
                        if (cvel(D, q, d) < 0) { cu = any ? cu - ul[d] : -ul[d]; any = true; }
                    }
                    const float a = __builtin_fmaf(ka, cu * cu, kc);
                    const float b = kb * cu;
                    float hq = a + b, hp = a - b;
                    if (addf) {
                        float cg = 0.f;
#pragma unroll
                        for (int d = 0; d < ldim(D); d++)
                            cg += float(cvel(D, q, d)) * float(B.g[d]);
                        const float fq = (B.fforce * float(kF16Scale) * cg) * rho;
                        hq += fq;
                        hp -= fq;
                    }
                    const float m = m_of(q);
                    store(q, __builtin_fmaf(m, f[q], hq));
                    store(p, __builtin_fmaf(m, f[p], hp));
                }
            } else {
                // opposite velocities share the even part of the
                // equilibrium: geq(q), geq(opp q) = A +- Bq with A = w (drho
                // + rho (4.5 cu^2 - 1.5 u^2)), Bq = 3 w rho cu, and the force
                // term changes sign
                store(0, om1 * f[0] + om * (float(wgt(D, 0)) * c0));
#pragma unroll
                for (int q = 1; q < Q; q++) {
                    const int p = opp(D, q);
                    if (p < q) continue;
                    float cu = 0.f;
#pragma unroll
                    for (int d = 0; d < ldim(D); d++) {
                        if (cvel(D, q, d) > 0) cu += ul[d];
                        if (cvel(D, q, d) < 0) cu -= ul[d];
                    }
                    const float wq = float(wgt(D, q));
                    const float a = wq * (c0 + rho * (4.5f * cu * cu));
                    const float bq = wq * rho * (3.f * cu);
                    float gq = om1 * f[q] + om * (a + bq);
                    float gp = om1 * f[p] + om * (a - bq);
                    if (addf) {
                        float cg = 0.f;
#pragma unroll
                        for (int d = 0; d < ldim(D); d++)
                            cg += float(cvel(D, q, d)) * float(B.g[d]);
                        const float fq = B.fforce * wq * rho * cg;
                        gq += fq;
                        gp -= fq;
                    }
                    store(q, gq);
                    store(p, gp);
                }
            }
        }
    }
}

// ---- the boundary / whole-lattice kernel --------------------------------------
// SCOPE_ALL: every node of every block, tiles of kTX x kTY nodes per z plane
// (initialisation and MATERIALIZE).  SCOPE_BND: the boundary segments only
// (the step's second kernel; the bulk kernel covers region 0).  Small
// cases run the whole step as one SCOPE_ALL launch (launch-bound there).
enum Scope : int { SCOPE_ALL = 0, SCOPE_BND = 1 };

template <int D, int ST, int MODE, int SCOPE>
__global__ __launch_bounds__(kTX * kTY) void k_lattice(KParams kp) {
    using T = typename Store<ST>::T;
    using E = typename Store<ST>::E;
    constexpr int Q = lq(D);
    constexpr bool EXACT = ST == MPLBHIP_STORE_F64;
    int b, x, y, z, wg;
    bool active;
    if constexpr (SCOPE == SCOPE_ALL) {
        wg = logical_wg(blockIdx.x, gridDim.x);
        b = wg_block(kp, kp.wg_block_all, wg);
        const BlockParams B = load_block(kp, b);
        const int t = wg - B.tile_start;
        const int t1 = fast_div(t, B.div_m[0], B.div_sh[0]);
        const int tx = t - t1 * B.tiles_x;
        z = fast_div(t1, B.div_m[1], B.div_sh[1]);
        const int ty = t1 - z * B.tiles_y;
        x = tx * kTX + threadIdx.x;
        y = ty * kTY + threadIdx.y;
        active = x < B.n[0] && y < B.n[1] && B.active;
        if constexpr (MODE == MODE_STEP || MODE == MODE_STEP_CHECK) {
            if (wg == B.tile_start && threadIdx.x == 0 && threadIdx.y == 0)
                kp.cur_out[b] = B.active ? 1 - load_parity(kp, b) : load_parity(kp, b);
        }
    } else {
        wg = blockIdx.x;
        const int v = kp.fold;
        b = wg_block(kp, kp.wg_block_bnd, wg);
        // the fold level indexes the block's arrays at run time: read the
        // three fields through scalar loads, not from a private copy of the
        // block (a dynamically indexed struct lives in scratch memory, and
        // every lane would write its 416-byte copy there)
        const MPLB_CONST BlockParams& Bc = const_view(kp.blocks)[b];
        const int i = (wg - Bc.bnd_wg_start[v]) * kBndWG + threadIdx.y * kTX + threadIdx.x;
        // segments start on workgroup boundaries: a lane past its
        // segment's end (padding) finds the next segment starting after it
        int s = Bc.seg_start[v];
        const int send = s + Bc.seg_count[v];
        while (s < send && i >= kp.segs[s].start + kp.segs[s].count) s++;
        active = s < send && i >= kp.segs[s].start && Bc.active;
        x = y = z = 0;
        if (active) {
            const BndSeg& S = kp.segs[s];
            int r = i - S.start;
            x = S.org[0] + r % S.ext[0];
            r /= S.ext[0];
            y = S.org[1] + r % S.ext[1];
            z = S.org[2] + r / S.ext[1];
        }
    }
    const BlockParams B = load_block(kp, b);
    // buffers of this block: steps read the latest fStage and write the
    // other; MATERIALIZE re-streams from the one before the latest
    const int cur = load_parity(kp, b);
    const int src = MODE == MODE_MATERIALIZE ? 1 - cur : cur;
    const void* fs_prev = kp.fs_buf[src];
    const void* per_prev = kp.per_buf[src];
    void* fs_next = kp.fs_buf[MODE >= MODE_INIT_FEQ ? cur : 1 - cur];
    void* per_next = kp.per_buf[MODE >= MODE_INIT_FEQ ? cur : 1 - cur];
    double acc[kNAcc];
#pragma unroll
    for (int a = 0; a < kNAcc; a++) acc[a] = 0.0;

    if (active) {
        const int32_t node = z * B.sz + y * B.sx + x;
        const int64_t mnode = B.mbase + ((int64_t)z * B.n[1] + y) * B.n[0] + x;
        const int r = region_code(x, B.n[0]) + 3 * region_code(y, B.n[1]) +
                      (D != 2 ? 9 * region_code(z, B.n[2]) : 0);
        const RegionProg* P = r != 0 ? &kp.progs[B.prog_base + r] : nullptr;
        const uint32_t om = obst_mask(kp, B, mnode);
        const bool solid = (P && P->kind == RK_SKIP) || (om & kObstSolid);
        const bool force_node = P ? P->force != 0 : true;
        T f[Q];  // EXACT: f; otherwise g = f - w (bulk) / f (boundary)
        bool full = false;  // f[] holds full values (boundary path)
        const E* prev = static_cast<const E*>(fs_prev) + B.base;
        E* next = static_cast<E*>(fs_next) + B.base;

        if (solid) {
            // ImmersedSolid: neither streamed nor collided; fStage unchanged,
            // f never set (zero), macros unchanged
            if constexpr (MODE != MODE_MATERIALIZE) {
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const int64_t i = (int64_t)q * B.sq + node;
                    next[i] = prev[i];
                }
            } else {
                if (kp.fhost && !kp.fhost_in) {
                    const int64_t nn = (int64_t)B.n[0] * B.n[1] * B.n[2];
                    const int64_t h = B.fbase + ((int64_t)z * B.n[1] + y) * B.n[0] + x;
#pragma unroll
                    for (int q = 0; q < Q; q++) kp.fhost[h + q * nn] = 0.0;
                }
            }
            solid_residual<D, MODE>(kp, mnode, acc);
            goto reduce;
        }
        if constexpr (MODE == MODE_INIT_FEQ || MODE == MODE_INIT_F) {
            // f^0 in double, from the macros (KerInitialiseBGK2nd[3D],
            // model_kernel.inc:193-219, 352-380) or from a host array
#pragma unroll
            for (int q = 0; q < Q; q++) {
                double fv;
                if constexpr (MODE == MODE_INIT_FEQ) {
                    const double* const* m0 = kp.macro0[0] ? kp.macro0 : kp.macro;
                    fv = feq_ref<D, double>(q, m0[0][mnode], m0[1][mnode], m0[2][mnode],
                                            D != 2 ? m0[3][mnode] : 0.0);
                } else {
                    const int64_t nn = (int64_t)B.n[0] * B.n[1] * B.n[2];
                    fv = kp.fhost[B.fbase + q * nn +
                                  ((int64_t)z * B.n[1] + y) * B.n[0] + x];
                }
                if constexpr (MODE == MODE_INIT_FEQ) {
                    if (kp.fhost) {
                        const int64_t nn = (int64_t)B.n[0] * B.n[1] * B.n[2];
                        kp.fhost[B.fbase + q * nn +
                                 ((int64_t)z * B.n[1] + y) * B.n[0] + x] = fv;
                    }
                }
                // the deviation is formed in double before rounding
                if constexpr (EXACT) f[q] = fv;
                else f[q] = T(fv - wgt(D, q));
                if (P && P->persist >= 0) {
                    T* per = static_cast<T*>(per_next);
                    per[persist_slot<D>(*P, x, y, z) * Q + q] = T(fv);
                }
            }
        } else if (MODE == MODE_MATERIALIZE && kp.fhost_in) {
            // the initial f, which no previous fStage can reproduce
            const int64_t nn = (int64_t)B.n[0] * B.n[1] * B.n[2];
            const int64_t h = B.fbase + ((int64_t)z * B.n[1] + y) * B.n[0] + x;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const double fv = kp.fhost[h + q * nn];
                if constexpr (EXACT) f[q] = fv;
                else f[q] = T(fv - wgt(D, q));
            }
        } else if (!P) {
            // bulk: every population streams in (KerStream3D Fluid branch)
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const E* p = (om >> q & 1u)
                                 ? at_u(prev, (int64_t)opp(D, q) * B.sq, (uint32_t)node)
                                 : at_u(prev, (int64_t)q * B.sq - pull_delta<D>(B, q),
                                        (uint32_t)node);
                if constexpr (ST == MPLBHIP_STORE_F16) f[q] = h2f_scaled(*p, kp.scale[q]);
                else f[q] = *p;
            }
        } else {
            // in the boundary kernel every active lane of a workgroup holds
            // the same region program (segments start on workgroup
            // boundaries): as a wave-uniform pointer its masks and ops are
            // scalar values and its per-population choices scalar branches
            if constexpr (SCOPE == SCOPE_BND) {
                const MPLB_CONST RegionProg& Pg = *const_view(first_lane_ptr(P));
                gather_masked<D, ST, MPLB_CONST RegionProg>(kp, B, Pg, x, y, z, f, fs_prev,
                                                            per_prev, om);
                boundary_program<D, ST, MPLB_CONST RegionProg>(kp, B, Pg, x, y, z, f,
                                                               fs_prev, per_prev);
            } else {
                gather_masked<D, ST>(kp, B, *P, x, y, z, f, fs_prev, per_prev, om);
                boundary_program<D, ST>(kp, B, *P, x, y, z, f, fs_prev, per_prev);
            }
            full = true;
            if constexpr (MODE != MODE_MATERIALIZE) {
                if (P->persist >= 0) {
                    T* per = static_cast<T*>(per_next);
                    const int64_t s = persist_slot<D>(*P, x, y, z);
#pragma unroll
                    for (int q = 0; q < Q; q++) per[s * Q + q] = f[q];
                }
            }
        }
        finish_node<D, ST, MODE>(kp, B, f, full, force_node, node, mnode, next, x, y, z, acc);
    }
reduce:
    if constexpr (MODE == MODE_STEP_CHECK || MODE == MODE_MATERIALIZE) {
        const int slot = SCOPE == SCOPE_BND ? kp.nbulk + wg : wg;
        reduce_store(acc, kp.partials + (int64_t)slot * kNAcc);
    }
}

// ---- the bulk kernel ------------------------------------------------------------
// Region 0 of every block: Fluid nodes whose Q populations all stream in
// (KerStream3D Fluid branch).  With FOLD it also runs the simple boundary
// regions (masked pull + EQM, nothing persisted) of the planes it covers,
// whose rows share cache lines with the interior; the other boundary
// regions stay in the boundary kernel, which keeps this one's register
// budget small.  Tiles of kTX x kTY nodes, z planes 1 .. n-2 (3D), or every
// z plane at fold level 3 (whose z-face planes then run their simple
// regions here too).  The block's first bulk workgroup also records the
// block's parity after the step.
template <int D, int ST, int MODE, bool OBST, int FOLD>
__global__ __launch_bounds__(kTX * kTY) void k_bulk(KParams kp) {
    using T = typename Store<ST>::T;
    using E = typename Store<ST>::E;
    constexpr int Q = lq(D);
    // fold level 3 spans every z plane (3D): its grid has its own tile starts
    constexpr bool ZALL = FOLD == 3 && D != 2;
    const int wg = kp.xcd_remap ? logical_wg(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int b = wg_block(kp, ZALL ? kp.wg_block_bulk3 : kp.wg_block_bulk, wg);
    const BlockParams B = load_block(kp, b);
    int t = wg - (ZALL ? B.bulk_tile_start3 : B.bulk_tile_start);
    // a block whose first tile sits on XCD 0 and whose tile count is a
    // multiple of 8 rows keeps each row's tiles on one XCD
    // (xgroup equals tiles_x: its divisor pair is tiles_x's)
    const int xg = ZALL ? B.xgroup3 : B.xgroup;
    if (kp.xgroup == 1 && xg) t = row_group_wg(t, xg, B.div_m[0], B.div_sh[0]);
    const bool first = t == 0;
    int x, y, z;
    if (kp.xgroup == 2 && B.rowwg) {
        // kTY consecutive x tiles of one row: the x-shifted pulls' straddled
        // lines are shared by waves of the same workgroup
        const int per_row = B.tiles_x / kTY;
        const int t1 = fast_div(t, B.div_m[3], B.div_sh[3]);
        const int tr = t - t1 * per_row;
        const int t2 = fast_div(t1, B.div_m[2], B.div_sh[2]);
        y = t1 - t2 * B.n[1];
        z = D != 2 ? (ZALL ? t2 : 1 + t2) : 0;
        x = (tr * kTY + (int)threadIdx.y) * kTX + threadIdx.x;
    } else {
        const int t1 = fast_div(t, B.div_m[0], B.div_sh[0]);
        const int tx = t - t1 * B.tiles_x;
        const int t2 = fast_div(t1, B.div_m[1], B.div_sh[1]);
        const int ty = t1 - t2 * B.tiles_y;
        z = D != 2 ? (ZALL ? t2 : 1 + t2) : 0;
        x = tx * kTX + threadIdx.x;
        y = ty * kTY + threadIdx.y;
    }
    const int cur = load_parity(kp, b);
    // a step reads the latest fStage and flips the block's parity;
    // MATERIALIZE re-streams from the fStage before the latest (as the
    // whole-lattice kernel does) and changes nothing
    if (MODE != MODE_MATERIALIZE && first && threadIdx.x == 0 && threadIdx.y == 0)
        kp.cur_out[b] = B.active ? 1 - cur : cur;
    const int src = MODE == MODE_MATERIALIZE ? 1 - cur : cur;
    double acc[kNAcc];
#pragma unroll
    for (int a = 0; a < kNAcc; a++) acc[a] = 0.0;
    // lanes: interior nodes, and (FOLD) the face and edge nodes of this
    // plane whose region program is simple (masked pull + EQM): their loads
    // share the wave's cache lines, which a separate kernel would fetch again
    const bool interior = x >= 1 && x < B.n[0] - 1 && y >= 1 && y < B.n[1] - 1 &&
                          (!ZALL || (z >= 1 && z < B.n[2] - 1));
    const RegionProg* P = nullptr;
    if (FOLD && !interior && x < B.n[0] && y < B.n[1]) {
        const bool yin = y >= 1 && y < B.n[1] - 1;
        if (FOLD >= 2 || yin) {
            P = &kp.progs[B.prog_base + region_code(x, B.n[0]) + 3 * region_code(y, B.n[1]) +
                          (ZALL ? 9 * region_code(z, B.n[2]) : 0)];
            if (!P->simple) P = nullptr;
        }
    }
    const bool active = B.active && (interior || P);
    if (active) {
        const int32_t node = z * B.sz + y * B.sx + x;
        const int64_t mnode = B.mbase + ((int64_t)z * B.n[1] + y) * B.n[0] + x;
        const E* prev = static_cast<const E*>(kp.fs_buf[src]) + B.base;
        E* next = static_cast<E*>(kp.fs_buf[1 - cur]) + B.base;
        uint32_t om = 0;
        if constexpr (OBST) om = obst_mask(kp, B, mnode);
        if (OBST && (om & kObstSolid)) {
            // obstacle node: nothing streams, collides or is read (its f is
            // zero when materialized)
            if constexpr (MODE == MODE_MATERIALIZE) {
                if (kp.fhost && !kp.fhost_in) {
                    const int64_t nn = (int64_t)B.n[0] * B.n[1] * B.n[2];
                    const int64_t h = B.fbase + ((int64_t)z * B.n[1] + y) * B.n[0] + x;
#pragma unroll
                    for (int q = 0; q < Q; q++) kp.fhost[h + q * nn] = 0.0;
                }
            }
            solid_residual<D, MODE>(kp, mnode, acc);
        } else {
            T f[Q];
            // every lane loads all Q populations, so the loads stay one
            // uniform burst; the pull mask applies afterwards.  A boundary
            // lane's unpulled source is moved back inside the block along
            // the axes where it left it: that row is one the wave (or its
            // workgroup) loads anyway, so no halo line is fetched for it.
            // FP32 and FP16 without obstacles: the populations as buffer
            // loads from prev - pull_shift, every pull offset a scalar
            constexpr bool kBufPull = ST != MPLBHIP_STORE_F64 && !OBST;
            const __amdgpu_buffer_rsrc_t rpull =
                buf_rsrc(reinterpret_cast<uint64_t>(prev) - (uint64_t)B.pull_shift * sizeof(E));
            // LEAN: fewer vector instructions per node, which bound the FP16
            // kernel (half the bytes per node of FP32) and the 2D FP32 one
            // (9 populations per node, boundary lanes in every wave of a
            // small block): 256 batched 2D cavities ran 10% faster.  3D FP32
            // and FP64, bound by HBM, keep the plain form: 3D FP32 ran 1-2%
            // slower on its fastest launch shapes with the lean one.
            constexpr bool LEAN =
                ST == MPLBHIP_STORE_F16 || (ST == MPLBHIP_STORE_F32 && D == 2);
            bool force_node = true;
            if constexpr (LEAN) {
                // one load per population and lane; for a lane in a simple
                // boundary region, sources across a face nothing streams
                // over are moved back inside the block (RegionProg::adj).
                // Waves without such a lane skip the shifts (a uniform
                // branch).
                auto load_all = [&](auto shifted) {
                    uint32_t sp[3] = {0u, 0u, 0u}, sm[3] = {0u, 0u, 0u};
                    if constexpr (decltype(shifted)::value) {
                        const uint32_t a = P ? P->adj : 0u;
                        sp[0] = (a & 1u) ? 1u : 0u;
                        sm[0] = (a & 2u) ? ~0u : 0u;
                        sp[1] = (a & 4u) ? (uint32_t)B.sx : 0u;
                        sm[1] = (a & 8u) ? (uint32_t)-B.sx : 0u;
                        if constexpr (D != 2) {
                            sp[2] = (a & 16u) ? (uint32_t)B.sz : 0u;
                            sm[2] = (a & 32u) ? (uint32_t)-B.sz : 0u;
                        }
                    }
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        uint32_t nd = (uint32_t)node;
                        if constexpr (decltype(shifted)::value) {
#pragma unroll
                            for (int d = 0; d < ldim(D); d++) {
                                if (cvel(D, q, d) > 0) nd += sp[d];
                                if (cvel(D, q, d) < 0) nd += sm[d];
                            }
                        }
                        E v;
                        if constexpr (kBufPull) {
                            v = buf_ld<E>(rpull, nd * (uint32_t)sizeof(E), B.pull_off[q]);
                        } else {
                            const bool bounce = OBST && (om >> q & 1u);
                            v = *(bounce
                                      ? at_u(prev, (int64_t)opp(D, q) * B.sq, (uint32_t)node)
                                      : at_u(prev, (int64_t)q * B.sq - pull_delta<D>(B, q), nd));
                        }
                        if constexpr (ST == MPLBHIP_STORE_F16) f[q] = h2f_scaled(v, kp.scale[q]);
                        else f[q] = v;
                    }
                };
                if (FOLD && __builtin_amdgcn_ballot_w64(P != nullptr) != 0)
                    load_all(std::true_type{});
                else
                    load_all(std::false_type{});
                if (FOLD && P) {
                    // the folded lanes of a wave (one wall face, or one row
                    // of it) share one or a few region programs: taken one
                    // program at a time, its ops are wave-uniform, read by
                    // scalar loads, and the per-population choices of the
                    // EQM application become scalar branches.  Every pass
                    // retires at least the first pending lane.  A population
                    // the program does not pull is written by one of its EQM
                    // applications before any reads it (a simple program
                    // keeps no state).  Deviations throughout, the
                    // wall-velocity terms from the host.
                    bool pending = true;
                    while (pending) {
                        const RegionProg* Pu = first_lane_ptr(P);
                        if (P == Pu) {
                            const MPLB_CONST RegionProg& Pc = *const_view(Pu);
                            force_node = Pc.force != 0;
                            const int nops = Pc.nops;
                            for (int o = 0; o < nops; o++) apply_eqm_dev<D>(f, Pc.op[o]);
                            pending = false;
                        }
                    }
                }
            } else {
                const uint32_t pull = P ? P->pull : ~0u;
                // every lane loads all Q populations, so the loads stay one
                // uniform burst; the pull mask applies afterwards.  A
                // boundary lane's unpulled source is moved back inside the
                // block along the axes where it left it: that row is one
                // the wave (or its workgroup) loads anyway, so no halo line
                // is fetched for it.
                int32_t adj[3] = {0, 0, 0};
                if (FOLD && P) {
                    adj[0] = x == 0 ? 1 : (x == B.n[0] - 1 ? -1 : 0);
                    adj[1] = y == 0 ? B.sx : (y == B.n[1] - 1 ? -B.sx : 0);
                    if (ZALL) adj[2] = z == 0 ? B.sz : (z == B.n[2] - 1 ? -B.sz : 0);
                }
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const bool bounce = OBST && (om >> q & 1u);
                    uint32_t nd = (uint32_t)node;
                    if (FOLD && P && !(pull >> q & 1u)) {
                        // only the components that point out of the block
                        if (cvel(D, q, 0) > 0 && adj[0] > 0) nd += 1;
                        if (cvel(D, q, 0) < 0 && adj[0] < 0) nd -= 1;
                        if (cvel(D, q, 1) > 0 && adj[1] > 0) nd += (uint32_t)B.sx;
                        if (cvel(D, q, 1) < 0 && adj[1] < 0) nd -= (uint32_t)B.sx;
                        if (ZALL && cvel(D, q, 2) > 0 && adj[2] > 0) nd += (uint32_t)B.sz;
                        if (ZALL && cvel(D, q, 2) < 0 && adj[2] < 0) nd -= (uint32_t)B.sz;
                    }
                    E v;
                    if constexpr (kBufPull) {
                        v = buf_ld<E>(rpull, nd * (uint32_t)sizeof(E), B.pull_off[q]);
                    } else {
                        v = *(bounce ? at_u(prev, (int64_t)opp(D, q) * B.sq, (uint32_t)node)
                                     : at_u(prev, (int64_t)q * B.sq - pull_delta<D>(B, q), nd));
                    }
                    f[q] = v;
                }
                if (FOLD && P) {
                    // full values; unpulled populations start at zero (the
                    // EQM applications set every one of them), then the
                    // boundary ops, one wave-uniform program at a time
#pragma unroll
                    for (int q = 0; q < Q; q++) {
                        if constexpr (ST != MPLBHIP_STORE_F64) f[q] += T(wgt(D, q));
                        f[q] = (pull >> q & 1u) ? f[q] : T(0);
                    }
                    bool pending = true;
                    while (pending) {
                        const RegionProg* Pu = first_lane_ptr(P);
                        if (P == Pu) {
                            const int nops = const_view(Pu)->nops;
                            for (int o = 0; o < nops; o++) {
                                BCOp op;
                                __builtin_memcpy(&op, (const BCOp*)&const_view(Pu)->op[o],
                                                 sizeof op);
                                apply_eqm<D, T>(f, op);
                            }
                            pending = false;
                        }
                    }
                }
                force_node = P ? P->force != 0 : true;
            }
            // buffer stores for FP32 only: 2-byte buffer stores ran 7%
            // slower than global ones (FP16 47.7k against 51.3k MLUPS)
            finish_node<D, ST, MODE, ST == MPLBHIP_STORE_F32>(
                kp, B, f, !LEAN && P != nullptr, force_node, node, mnode, next, x, y, z, acc);
        }
    }
    if constexpr (MODE == MODE_STEP_CHECK || MODE == MODE_MATERIALIZE)
        reduce_store(acc, kp.partials + (int64_t)wg * kNAcc);
}

// ---- warm-start initial state ---------------------------------------------------
// f = feq(rho, u) + f1, with f1 the first-order (Chapman-Enskog)
// non-equilibrium part of BGK in lattice units (velocity u_L = u / CS,
// unit node spacing and time step):
//   f1_q = -3 tau_L w_q rho sum_ab (c_qa c_qb - delta_ab / 3) d_a u_L,b ,
// tau_L = tau / dt + 1/2 of the block.  Gradients are central differences,
// one-sided where a neighbour is outside the block or an obstacle node.
// Obstacle nodes get f = 0 (ImmersedSolid).  Writes the host-layout array
// fhost that MODE_INIT_F then collides.  No reference counterpart.
template <int D>
__global__ void k_lift(KParams kp, int64_t nodes) {
    constexpr int Q = lq(D);
    const double* const* m0 = kp.macro0[0] ? kp.macro0 : kp.macro;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nodes;
         i += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = kp.nblocks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (kp.blocks[mid].mbase <= i) lo = mid; else hi = mid - 1;
        }
        const BlockParams& B = kp.blocks[lo];
        const int64_t local = i - B.mbase;
        const int64_t nn = (int64_t)B.n[0] * B.n[1] * B.n[2];
        double* out = kp.fhost + B.fbase + local;
        if (obst_mask(kp, B, i) & kObstSolid) {
            for (int q = 0; q < Q; q++) out[q * nn] = 0.0;
            continue;
        }
        const int c[3] = {(int)(local % B.n[0]), (int)((local / B.n[0]) % B.n[1]),
                          (int)(local / ((int64_t)B.n[0] * B.n[1]))};
        const int64_t stride[3] = {1, B.n[0], (int64_t)B.n[0] * B.n[1]};
        double g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // g[a][b] = d_a u_b
        for (int a = 0; a < ldim(D); a++) {
            const bool hp = c[a] + 1 < B.n[a] &&
                            !(obst_mask(kp, B, i + stride[a]) & kObstSolid);
            const bool hm = c[a] > 0 && !(obst_mask(kp, B, i - stride[a]) & kObstSolid);
            for (int b = 0; b < ldim(D); b++) {
                const double* u = m0[1 + b];
                const double u0 = u[i];
                double d = 0.0;
                if (hp && hm) d = 0.5 * (u[i + stride[a]] - u[i - stride[a]]);
                else if (hp) d = u[i + stride[a]] - u0;
                else if (hm) d = u0 - u[i - stride[a]];
                g[a][b] = d / kCS;
            }
        }
        const double rho = m0[0][i];
        const double tl = B.tau / B.dt + 0.5;
        const double div = g[0][0] + g[1][1] + g[2][2];
        for (int q = 0; q < Q; q++) {
            double s = 0.0;
            for (int a = 0; a < ldim(D); a++)
                for (int b = 0; b < ldim(D); b++) s += cvel(D, q, a) * cvel(D, q, b) * g[a][b];
            const double f1 = -3.0 * tl * wgt(D, q) * rho * (s - div / 3.0);
            out[q * nn] = feq_ref<D, double>(q, rho, m0[1][i], m0[2][i],
                                             D != 2 ? m0[3][i] : 0.0) + f1;
        }
    }
}

hipError_t launch_lift(int dim, const KParams& kp, int64_t nodes, hipStream_t s) {
    int64_t grid = (nodes + 255) / 256;
    if (grid > 16384) grid = 16384;
    if (grid < 1) grid = 1;
    if (dim == 2) hipLaunchKernelGGL(k_lift<2>, dim3((unsigned)grid), dim3(256), 0, s, kp, nodes);
    else hipLaunchKernelGGL(k_lift<3>, dim3((unsigned)grid), dim3(256), 0, s, kp, nodes);
    return hipGetLastError();
}

// ---- FP16 storage: two nodes per lane ---------------------------------------------
// The BGK update of one node in deviation form, FP32 arithmetic, as in
// finish_node's step path: g in (deviations), h out (storage units).
template <int D>
__device__ __forceinline__ void collide_dev(const KParams& kp, const BlockParams& B,
                                            const float* g, float* h) {
    constexpr int Q = lq(D);
    float drho = 0.f, j[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < Q; q++) {
        drho += g[q];
#pragma unroll
        for (int d = 0; d < ldim(D); d++) {
            if (cvel(D, q, d) > 0) j[d] += g[q];
            if (cvel(D, q, d) < 0) j[d] -= g[q];
        }
    }
    const float rho = 1.f + drho;
    const float inv = __builtin_amdgcn_rcpf(rho);  // v_rcp_f32, 1 ulp
    float ul[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < ldim(D); d++) {
        ul[d] = j[d] * inv;
        if (kp.force_macros) ul[d] += B.fhalf_dtg[d];
    }
    const float usq = ul[0] * ul[0] + ul[1] * ul[1] + ul[2] * ul[2];
    const float om = B.fomega, om1 = B.fone_minus_omega;
    const bool addf = kp.force_type == 1;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        float cu = 0.f;
#pragma unroll
        for (int d = 0; d < ldim(D); d++) {
            if (cvel(D, q, d) > 0) cu += ul[d];
            if (cvel(D, q, d) < 0) cu -= ul[d];
        }
        const float wq = float(wgt(D, q));
        const float geq = wq * (drho + rho * (3.f * cu + 4.5f * cu * cu - 1.5f * usq));
        float gn = om1 * g[q] + om * geq;
        if (addf) {
            float cg = 0.f;
#pragma unroll
            for (int d = 0; d < ldim(D); d++) cg += float(cvel(D, q, d)) * float(B.g[d]);
            gn += B.fforce * wq * rho * cg;
        }
        h[q] = gn * kp.inv_scale[q];
    }
}

__device__ __forceinline__ uint32_t ld_u32(const __half* p) {
    return *reinterpret_cast<const uint32_t*>(p);
}
__device__ __forceinline__ float h_lo(uint32_t v) {
    return __half2float(__ushort_as_half((unsigned short)(v & 0xFFFFu)));
}
__device__ __forceinline__ float h_hi(uint32_t v) {
    return __half2float(__ushort_as_half((unsigned short)(v >> 16)));
}

// FP16 bulk step with two neighbouring nodes (x0, x0 + 1; x0 even) per
// lane: every population moves as one 4-byte half pair per lane.  A pull
// along x by one node crosses the pair boundary; the missing half comes
// from the neighbouring lane (shuffle), and the lanes at the wave's ends
// load it themselves.  Interior nodes only: the boundary kernel runs every
// boundary region (fold level 0).  Rows of 128 nodes per wave.
template <int D>
__global__ __launch_bounds__(kTX * kTY) void k_bulk_h2(KParams kp) {
    constexpr int Q = lq(D);
    const int wg = (int)blockIdx.x;
    const int b = wg_block(kp, kp.wg_block_h2, wg);
    const BlockParams B = load_block(kp, b);
    int t = wg - B.h2_tile_start;
    const bool first = t == 0;
    const int tx = t % B.tiles_x2;
    t /= B.tiles_x2;
    const int ty = t % B.tiles_y;
    const int z = D != 2 ? 1 + t / B.tiles_y : 0;
    const int lane = threadIdx.x;
    const int x0 = tx * 2 * kTX + 2 * lane;
    const int y = ty * kTY + threadIdx.y;
    const int cur = load_parity(kp, b);
    if (first && lane == 0 && threadIdx.y == 0) kp.cur_out[b] = B.active ? 1 - cur : cur;
    if (!B.active) return;
    const bool yin = y >= 1 && y < B.n[1] - 1;
    const bool a0 = yin && x0 >= 1 && x0 < B.n[0] - 1;
    const bool a1 = yin && x0 + 1 >= 1 && x0 + 1 < B.n[0] - 1;
    // every lane loads (the shuffles need all of them).  FP16 rows are
    // padded so that x0 - 1 .. x0 + 2 of every lane of the grid lie in the
    // row; padding rows load the last interior row instead.  An active node
    // only ever uses x - 1 .. x + 1 of its own row.
    const int yc = y < B.n[1] - 1 ? y : B.n[1] - 1;
    const int32_t node = z * B.sz + yc * B.sx + x0;
    const __half* prev = static_cast<const __half*>(kp.fs_buf[cur]) + B.base;
    __half* next = static_cast<__half*>(kp.fs_buf[1 - cur]) + B.base;
    float g0[Q], g1[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int cx = cvel(D, q, 0);
        // the aligned pair of the source row, before the x shift
        const int32_t dyz = cvel(D, q, 1) * B.sx + (D != 2 ? cvel(D, q, 2) * B.sz : 0);
        const __half* row = prev + (int64_t)q * B.sq - dyz;
        const uint32_t P = ld_u32(row + node);
        float lo, hi;
        if (cx == 0) {
            lo = h_lo(P);
            hi = h_hi(P);
        } else if (cx > 0) {
            // (x0 - 1, x0): the low half of the pair is the left lane's high
            const uint32_t L = (uint32_t)__shfl_up((int)P, 1, 64);
            const float left = lane == 0 ? __half2float(row[node - 1]) : h_hi(L);
            lo = left;
            hi = h_lo(P);
        } else {
            // (x0 + 1, x0 + 2): the high half is the right lane's low
            const uint32_t R = (uint32_t)__shfl_down((int)P, 1, 64);
            const float right = lane == 63 ? __half2float(row[node + 2]) : h_lo(R);
            lo = h_hi(P);
            hi = right;
        }
        g0[q] = lo * kp.scale[q];
        g1[q] = hi * kp.scale[q];
    }
    if (!(a0 || a1)) return;
    float h0[Q], h1[Q];
    collide_dev<D>(kp, B, g0, h0);
    collide_dev<D>(kp, B, g1, h1);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        __half* p = next + (int64_t)q * B.sq + node;
        if (a0 && a1) {
            const uint32_t v = (uint32_t)__half_as_ushort(__float2half_rn(h0[q])) |
                               ((uint32_t)__half_as_ushort(__float2half_rn(h1[q])) << 16);
            __builtin_nontemporal_store(v, reinterpret_cast<uint32_t*>(p));
        } else if (a0) {
            *p = __float2half_rn(h0[q]);
        } else {
            p[1] = __float2half_rn(h1[q]);
        }
    }
}

hipError_t launch_bulk_h2(int dim, const KParams& kp, hipStream_t s) {
    if (kp.nbulk2 <= 0) return hipSuccess;
    dim3 blk(kTX, kTY, 1);
    if (dim == 2) hipLaunchKernelGGL(k_bulk_h2<2>, dim3(kp.nbulk2), blk, 0, s, kp);
    else hipLaunchKernelGGL(k_bulk_h2<3>, dim3(kp.nbulk2), blk, 0, s, kp);
    return hipGetLastError();
}

// ---- restart of an instance at rest (batched solver fallback) --------------------
// The macros of block b and their residual copies become rho = 1, u = 0, as
// set_macros leaves them before a rest start; an INIT_FEQ launch in which
// only the restarting blocks are active then forms their fStage from them.
__global__ void k_rest_macros(KParams kp, int b) {
    const BlockParams& B = kp.blocks[b];
    const int64_t nn = (int64_t)B.n[0] * B.n[1] * B.n[2];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nn;
         i += (int64_t)gridDim.x * blockDim.x) {
        for (int v = 0; v < 4; v++) {
            kp.macro[v][B.mbase + i] = v == 0 ? 1.0 : 0.0;
            kp.macro_copy[v][B.mbase + i] = v == 0 ? 1.0 : 0.0;
        }
    }
}

hipError_t launch_rest_macros(const KParams& kp, int b, int64_t nodes, hipStream_t s) {
    int64_t grid = (nodes + 255) / 256;
    if (grid > 4096) grid = 4096;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_rest_macros, dim3((unsigned)grid), dim3(256), 0, s, kp, b);
    return hipGetLastError();
}

// fStage halo slab copy of one block connection (Field::TransferHalos ->
// ops_halo_transfer, field.h:94-99, slabs of field.h:215-395).  One thread
// per slab cell and population (grid.y = population), 32-bit cell
// arithmetic: the slab is one layer of a face.  A target halo cell is read
// by the target block's pulls only for the populations that stream inward
// across the face, so only those are copied, except for cells another
// connection reads (flag bit 1).  Across a face the slab has no contiguous
// axis, so every copied element costs a cache line each way.
template <typename E>
__global__ __launch_bounds__(256) void k_halo(HaloParams hp) {
    const int ncell = hp.ext[0] * hp.ext[1] * hp.ext[2];
    const int cell = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (cell >= ncell) return;
    const int q = (int)blockIdx.y;
    const uint8_t fl = hp.other_flag ? hp.other_flag[cell] : 0;
    if (!((hp.inward >> q) & 1u) && !(fl & 2)) return;
    const int a = cell % hp.ext[0];
    const int rest = cell / hp.ext[0];
    const int bb = rest % hp.ext[1];
    const int cc = rest / hp.ext[1];
    const E* src = static_cast<const E*>((fl & 1) ? hp.src_other : hp.src_cur);
    const int64_t si = hp.src_base + (int64_t)q * hp.src_sq +
                       (int64_t)(hp.org_src[2] + cc) * hp.src_sz +
                       (int64_t)(hp.org_src[1] + bb) * hp.src_sx + hp.org_src[0] + a;
    const int64_t di = hp.dst_base + (int64_t)hp.qmap[q] * hp.dst_sq + hp.dst_org +
                       a * hp.dstep[0] + bb * hp.dstep[1] + cc * hp.dstep[2];
    static_cast<E*>(hp.dst)[di] = src[si];
}

// Several block connections in one launch: workgroup w belongs to the link
// whose range [wg_start[l], wg_start[l+1]) holds it; within a link the work
// is k_halo's.  The host merges only links that neither read a halo cell
// another merged link writes in this exchange nor write a cell another
// link wrote before them, and gives each cell the target populations a
// later merged link writes there (sup), so no two threads write one element
// and every read sees what the one-link-at-a-time order would show.
template <typename E>
__global__ __launch_bounds__(256) void k_halo_multi(const HaloParams* __restrict__ links,
                                                    const int32_t* __restrict__ wg_start,
                                                    int nlinks) {
    const int wg = (int)blockIdx.x;
    int lo = 0, hi = nlinks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (wg_start[mid] <= wg) lo = mid; else hi = mid - 1;
    }
    const HaloParams& hp = links[lo];
    const int ncell = hp.ext[0] * hp.ext[1] * hp.ext[2];
    const int cell = (wg - wg_start[lo]) * 256 + (int)threadIdx.x;
    if (cell >= ncell) return;
    const int q = (int)blockIdx.y;
    if (q >= hp.Q) return;
    const uint8_t fl = hp.other_flag ? hp.other_flag[cell] : 0;
    if (!((hp.inward >> q) & 1u) && !(fl & 2)) return;
    if (hp.sup && ((hp.sup[cell] >> hp.qmap[q]) & 1u)) return;
    const int a = cell % hp.ext[0];
    const int rest = cell / hp.ext[0];
    const int bb = rest % hp.ext[1];
    const int cc = rest / hp.ext[1];
    const E* src = static_cast<const E*>((fl & 1) ? hp.src_other : hp.src_cur);
    const int64_t si = hp.src_base + (int64_t)q * hp.src_sq +
                       (int64_t)(hp.org_src[2] + cc) * hp.src_sz +
                       (int64_t)(hp.org_src[1] + bb) * hp.src_sx + hp.org_src[0] + a;
    const int64_t di = hp.dst_base + (int64_t)hp.qmap[q] * hp.dst_sq + hp.dst_org +
                       a * hp.dstep[0] + bb * hp.dstep[1] + cc * hp.dstep[2];
    static_cast<E*>(hp.dst)[di] = src[si];
}

// Per-block sums of the per-workgroup partials, one workgroup per block:
// the threads stride over the block's slots (its bulk tiles, then its
// boundary workgroups; or its whole-lattice tiles), then the waves' sums
// (shuffles) and the workgroup's (LDS) are added; every order is fixed, so
// the sums are reproducible.
constexpr int kRedThreads = 1024;
__global__ __launch_bounds__(kRedThreads) void k_reduce_blocks(
    const double* __restrict__ partials, const BlockParams* __restrict__ blocks, int nblocks,
    int ntiles, int nbulk, int nbnd, int split, int fold, double* __restrict__ out) {
    const int b = blockIdx.x;
    const bool last = b + 1 == nblocks;
    int lo0, hi0, lo1 = 0, hi1 = 0;
    if (split) {
        // the fold-3 bulk grid (every z plane) has its own tile starts
        const bool zall = fold == 3;
        lo0 = zall ? blocks[b].bulk_tile_start3 : blocks[b].bulk_tile_start;
        hi0 = last ? nbulk : (zall ? blocks[b + 1].bulk_tile_start3 : blocks[b + 1].bulk_tile_start);
        lo1 = nbulk + blocks[b].bnd_wg_start[fold];
        hi1 = nbulk + (last ? nbnd : blocks[b + 1].bnd_wg_start[fold]);
    } else {
        lo0 = blocks[b].tile_start;
        hi0 = last ? ntiles : blocks[b + 1].tile_start;
    }
    const int n0 = hi0 - lo0, count = n0 + (hi1 - lo1);
    // one pass over the slots, all kNAcc sums per thread (a slot's kNAcc
    // doubles are contiguous), then waves and workgroup in a fixed order
    double s[kNAcc];
#pragma unroll
    for (int a = 0; a < kNAcc; a++) s[a] = 0.0;
    for (int i = threadIdx.x; i < count; i += kRedThreads) {
        const int slot = i < n0 ? lo0 + i : lo1 + (i - n0);
        const double* p = partials + (int64_t)slot * kNAcc;
#pragma unroll
        for (int a = 0; a < kNAcc; a++) s[a] += p[a];
    }
    constexpr int kWaves = kRedThreads / 64;
    __shared__ double red[kWaves][kNAcc];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < kNAcc; a++) {
        const double w = wave_sum(s[a]);
        if (lane == 0) red[wv][a] = w;
    }
    __syncthreads();
    if (threadIdx.x < kNAcc) {
        double t = 0.0;
        for (int k = 0; k < kWaves; k++) t += red[k][threadIdx.x];
        out[(int64_t)b * kNAcc + threadIdx.x] = t;
    }
}

// the total over the blocks, in the same fixed-order fashion
__global__ __launch_bounds__(256) void k_reduce_total(double* __restrict__ out, int nblocks) {
    __shared__ double red[256];
    for (int a = 0; a < kNAcc; a++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += 256) s += out[(int64_t)b * kNAcc + a];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[(int64_t)nblocks * kNAcc + a] = red[0];
        __syncthreads();
    }
}

__global__ void k_fill_bytes(uint32_t* p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---- host-side launch helpers ----------------------------------------------------
// part: 0 = the whole step (bulk then boundary), 1 = bulk only, 2 = boundary
// only (timed runs bracket the two separately); other modes ignore it
template <int D, int ST, int MODE>
static void launch_step(const KParams& kp, int part, hipStream_t s) {
    dim3 blk(kTX, kTY, 1);
    // the two-node FP16 kernel exists for D2Q9 and D3Q19
    if (ST == MPLBHIP_STORE_F16 && MODE == MODE_STEP && D != 15 && kp.h2 && !kp.fused) {
        if (part != 2) (void)launch_bulk_h2(D, kp, s);
        if (part != 1 && kp.nbnd > 0)
            hipLaunchKernelGGL((k_lattice<D, ST, MODE, SCOPE_BND>), dim3(kp.nbnd), blk, 0, s, kp);
        return;
    }
    if (kp.fused) {
        if (part != 2)
            hipLaunchKernelGGL((k_lattice<D, ST, MODE, SCOPE_ALL>), dim3(kp.ntiles), blk, 0, s,
                               kp);
        return;
    }
    if (part != 2 && kp.nbulk > 0) {
        const dim3 g(kp.nbulk);
        if (kp.obst) {
            if (kp.fold == 3 && D != 2)
                hipLaunchKernelGGL((k_bulk<D, ST, MODE, true, 3>), g, blk, 0, s, kp);
            else if (kp.fold >= 2)
                hipLaunchKernelGGL((k_bulk<D, ST, MODE, true, 2>), g, blk, 0, s, kp);
            else if (kp.fold == 1)
                hipLaunchKernelGGL((k_bulk<D, ST, MODE, true, 1>), g, blk, 0, s, kp);
            else hipLaunchKernelGGL((k_bulk<D, ST, MODE, true, 0>), g, blk, 0, s, kp);
        } else {
            if (kp.fold == 3 && D != 2)
                hipLaunchKernelGGL((k_bulk<D, ST, MODE, false, 3>), g, blk, 0, s, kp);
            else if (kp.fold >= 2)
                hipLaunchKernelGGL((k_bulk<D, ST, MODE, false, 2>), g, blk, 0, s, kp);
            else if (kp.fold == 1)
                hipLaunchKernelGGL((k_bulk<D, ST, MODE, false, 1>), g, blk, 0, s, kp);
            else hipLaunchKernelGGL((k_bulk<D, ST, MODE, false, 0>), g, blk, 0, s, kp);
        }
    }
    if (part != 1 && kp.nbnd > 0)
        hipLaunchKernelGGL((k_lattice<D, ST, MODE, SCOPE_BND>), dim3(kp.nbnd), blk, 0, s, kp);
}

template <int D, int ST>
static hipError_t launch_mode(int mode, int part, const KParams& kp, hipStream_t s) {
    dim3 blk(kTX, kTY, 1), grd(kp.ntiles, 1, 1);
    switch (mode) {
        case MODE_STEP:
            launch_step<D, ST, MODE_STEP>(kp, part, s);
            break;
        case MODE_STEP_CHECK:
            launch_step<D, ST, MODE_STEP_CHECK>(kp, part, s);
            break;
        case MODE_MATERIALIZE:
            hipLaunchKernelGGL((k_lattice<D, ST, MODE_MATERIALIZE, SCOPE_ALL>), grd, blk, 0, s, kp);
            break;
        case MODE_INIT_FEQ:
            hipLaunchKernelGGL((k_lattice<D, ST, MODE_INIT_FEQ, SCOPE_ALL>), grd, blk, 0, s, kp);
            break;
        case MODE_INIT_F:
            hipLaunchKernelGGL((k_lattice<D, ST, MODE_INIT_F, SCOPE_ALL>), grd, blk, 0, s, kp);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// MATERIALIZE through the step's bulk and boundary kernels (their launch
// shape and partial-sum slots); the whole-lattice launch remains for t = 0
// (f^0 from a host array) and fused cases
hipError_t launch_materialize_split(int dim, int storage, const KParams& kp, hipStream_t s) {
    if (kp.fused || kp.fhost_in || (dim != 2 && dim != 3 && dim != 15)) return hipErrorInvalidValue;
    if (dim == 15) {
        if (storage == MPLBHIP_STORE_F64) launch_step<15, MPLBHIP_STORE_F64, MODE_MATERIALIZE>(kp, 0, s);
        else if (storage == MPLBHIP_STORE_F32) launch_step<15, MPLBHIP_STORE_F32, MODE_MATERIALIZE>(kp, 0, s);
        else launch_step<15, MPLBHIP_STORE_F16, MODE_MATERIALIZE>(kp, 0, s);
    } else if (dim == 2) {
        if (storage == MPLBHIP_STORE_F64) launch_step<2, MPLBHIP_STORE_F64, MODE_MATERIALIZE>(kp, 0, s);
        else if (storage == MPLBHIP_STORE_F32) launch_step<2, MPLBHIP_STORE_F32, MODE_MATERIALIZE>(kp, 0, s);
        else launch_step<2, MPLBHIP_STORE_F16, MODE_MATERIALIZE>(kp, 0, s);
    } else {
        if (storage == MPLBHIP_STORE_F64) launch_step<3, MPLBHIP_STORE_F64, MODE_MATERIALIZE>(kp, 0, s);
        else if (storage == MPLBHIP_STORE_F32) launch_step<3, MPLBHIP_STORE_F32, MODE_MATERIALIZE>(kp, 0, s);
        else launch_step<3, MPLBHIP_STORE_F16, MODE_MATERIALIZE>(kp, 0, s);
    }
    return hipGetLastError();
}

// dim: the lattice key, 2 (D2Q9), 3 (D3Q19) or 15 (D3Q15)
hipError_t launch_lattice(int dim, int storage, int mode, const KParams& kp, hipStream_t s,
                          int part) {
    if (dim == 15) {
        if (storage == MPLBHIP_STORE_F64) return launch_mode<15, MPLBHIP_STORE_F64>(mode, part, kp, s);
        if (storage == MPLBHIP_STORE_F32) return launch_mode<15, MPLBHIP_STORE_F32>(mode, part, kp, s);
        return launch_mode<15, MPLBHIP_STORE_F16>(mode, part, kp, s);
    }
    if (dim == 2) {
        if (storage == MPLBHIP_STORE_F64) return launch_mode<2, MPLBHIP_STORE_F64>(mode, part, kp, s);
        if (storage == MPLBHIP_STORE_F32) return launch_mode<2, MPLBHIP_STORE_F32>(mode, part, kp, s);
        return launch_mode<2, MPLBHIP_STORE_F16>(mode, part, kp, s);
    }
    if (storage == MPLBHIP_STORE_F64) return launch_mode<3, MPLBHIP_STORE_F64>(mode, part, kp, s);
    if (storage == MPLBHIP_STORE_F32) return launch_mode<3, MPLBHIP_STORE_F32>(mode, part, kp, s);
    return launch_mode<3, MPLBHIP_STORE_F16>(mode, part, kp, s);
}

hipError_t launch_halo(const HaloParams& hp, hipStream_t s) {
    const int64_t ncell = (int64_t)hp.ext[0] * hp.ext[1] * hp.ext[2];
    if (ncell >= (int64_t)1 << 31 || hp.Q < 1 || hp.Q > kMaxQ) return hipErrorInvalidValue;
    if (ncell == 0) return hipSuccess;
    const dim3 grid((unsigned)((ncell + 255) / 256), (unsigned)hp.Q);
    if (hp.elem_bytes == 8)
        hipLaunchKernelGGL(k_halo<double>, grid, dim3(256), 0, s, hp);
    else if (hp.elem_bytes == 4)
        hipLaunchKernelGGL(k_halo<float>, grid, dim3(256), 0, s, hp);
    else
        hipLaunchKernelGGL(k_halo<__half>, grid, dim3(256), 0, s, hp);
    return hipGetLastError();
}

hipError_t launch_halo_multi(const HaloParams* links, const int32_t* wg_start, int nlinks,
                             int nwg, int Q, int elem_bytes, hipStream_t s) {
    if (nlinks < 1 || nwg < 1 || Q < 1 || Q > kMaxQ) return hipErrorInvalidValue;
    const dim3 grid((unsigned)nwg, (unsigned)Q);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(k_halo_multi<double>, grid, dim3(256), 0, s, links, wg_start, nlinks);
    else if (elem_bytes == 4)
        hipLaunchKernelGGL(k_halo_multi<float>, grid, dim3(256), 0, s, links, wg_start, nlinks);
    else
        hipLaunchKernelGGL(k_halo_multi<__half>, grid, dim3(256), 0, s, links, wg_start, nlinks);
    return hipGetLastError();
}

hipError_t launch_reduce(const double* partials, const BlockParams* blocks, int nblocks,
                         int ntiles, int nbulk, int nbnd, int split, int fold, double* out,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_blocks, dim3(nblocks), dim3(kRedThreads), 0, s, partials, blocks,
                       nblocks, ntiles, nbulk, nbnd, split, fold, out);
    hipLaunchKernelGGL(k_reduce_total, dim3(1), dim3(256), 0, s, out, nblocks);
    return hipGetLastError();
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy16(const u32x4* __restrict__ src,
                                                u32x4* __restrict__ dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        __builtin_nontemporal_store(src[i], dst + i);
}

hipError_t launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t s) {
    // one 16-byte element per thread up to 16 GiB (a grid-stride loop
    // over 8192 workgroups copied 2 GiB at 4.75 TB/s, this form at 6.28)
    const int64_t n = bytes / 16;
    const int64_t grid = std::min<int64_t>((n + 255) / 256, int64_t(1) << 22);
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)grid), dim3(256), 0, s,
                       static_cast<const u32x4*>(src), static_cast<u32x4*>(dst), n);
    return hipGetLastError();
}

// ---- scratch of the allocation-free entry point --------------------------------
// Per-workgroup partial sums and per-block totals of the FP64 reductions of
// mplbhip_mass_noalloc live in this file-scope array instead of memory the
// library allocates.  Each GPU the library runs on holds its own copy until
// the process ends; every call on one GPU shares it (calls that overlap on
// one GPU race).  Host code never passes the array's name to a kernel (on
// the host it is only a stand-in): it takes the device address once per
// device from hipGetSymbolAddress, with that device current.
static __device__ __align__(256) double scratch[kScratchDoubles];

hipError_t scratch_address(double** ptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    *ptr = e == hipSuccess ? static_cast<double*>(p) : nullptr;
    return e;
}

hipError_t launch_fill(void* p, int64_t bytes, hipStream_t s) {
    const int64_t n = bytes / 4;
    int grid = (int)((n + 255) / 256);
    if (grid > 8192) grid = 8192;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_fill_bytes, dim3(grid), dim3(256), 0, s,
                       static_cast<uint32_t*>(p), n, 0u);
    return hipGetLastError();
}

}  // namespace mplbhip
